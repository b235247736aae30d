"""Symmetric Verlet lists and ghost_put<add_>, CPU checks of the oracle, by the
property the reference's vector_dist_symmetric_cell_list checks
(Vector/tests/vector_dist_cell_list_tests.cpp:443-649): a pass over the
symmetric list that adds each pair to both ends, followed by ghost_put<add_>,
gives every domain particle the neighbour count and the neighbour id set of the
full list (brute force over domain + ghosts), and the LJ forces of the full
list.  The rule must hold for any total order of the particles: identity and a
random permutation are both checked."""
import numpy as np
import pytest

from oracle import oracle as ora
from openfpm_pdata_amd import workloads


def snap(wl):
    """Positions on multiples of 2^-20 in [0, 1): x +- 1 is exact in FP32 and FP64,
    so a ghost pair's distance is the same from both ends (no cut-off ties)."""
    q = np.round(wl.pos.astype(np.float64) * 2.0 ** 20) / 2.0 ** 20
    wl.pos[...] = np.where(q >= 1.0, q - 1.0, q).astype(wl.pos.dtype)
    return wl


def _setup(dt, n=3000, seed=1):
    wl = snap(workloads.uniform_lj(n=n, dtype=dt, seed=seed))
    pos = np.ascontiguousarray(wl.pos.astype(dt))
    g = wl.ghost
    blo, bhi, sv = ora.local_ghost_boxes(wl.box_lo, wl.box_hi, [-g] * 3, [g] * 3, wl.bc, dt)
    shifts = ora.shift_vectors(wl.box_lo, wl.box_hi, dt)
    gp, opid, osid = ora.ghost_emit(pos, wl.n, blo, bhi, sv, shifts)
    allpos = np.ascontiguousarray(np.concatenate([pos, gp], axis=1))
    off, nbr = ora.brute_nn(allpos, wl.n, wl.r_cut)
    return wl, allpos, opid.astype(np.int64), osid.astype(np.int64), off.astype(np.int64), nbr.astype(np.int64)


def _lj(allpos, p, q, sigma, eps):
    r = (allpos[:, p] - allpos[:, q]).astype(np.float64)
    d2 = (r ** 2).sum(axis=0)
    s6 = (sigma ** 2 / d2) ** 3
    f = 24.0 * eps * (2.0 * s6 * s6 - s6) / d2
    return f * r


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["identity", "random"])
def test_sym_list_then_put_gives_full_counts_ids_and_forces(dt, order):
    wl, allpos, opid, osid, off, nbr = _setup(dt)
    n, n_all = wl.n, allpos.shape[1]
    key = np.arange(n) if order == "identity" else np.random.default_rng(3).permutation(n)
    off_s, nbr_s = ora.verlet_sym(off, nbr, key, np.concatenate([key, np.zeros(n_all - n, np.int64)]), n, opid,
                                  osid)
    rows_full = np.repeat(np.arange(n), np.diff(off))
    rows_sym = np.repeat(np.arange(n), np.diff(off_s))
    # no self entries; about half the full list's non-self entries
    assert not np.any(nbr_s == rows_sym)
    assert len(nbr_s) * 2 == len(nbr) - n
    # counts: each kept pair adds one to both ends, then ghost_put<add_>
    c = np.bincount(rows_sym, minlength=n_all) + np.bincount(nbr_s, minlength=n_all)
    c = ora.ghost_put_add(c, n, opid)
    assert np.array_equal(c[:n], np.diff(off) - 1)
    # neighbour id sets (ghost_put<merge_> of the ids, by source particle)
    src = lambda q: np.where(q >= n, opid[np.maximum(q - n, 0)], q)
    full = sorted(zip(rows_full[nbr != rows_full].tolist(), src(nbr[nbr != rows_full]).tolist()))
    half = list(zip(rows_sym.tolist(), src(nbr_s).tolist())) + list(zip(src(nbr_s).tolist(), rows_sym.tolist()))
    assert sorted(half) == full
    # LJ forces, Newton's third law on the symmetric list, against the full list
    sig, eps = wl.params["sigma"], wl.params["eps"]
    F = np.zeros((3, n_all))
    f = _lj(allpos, rows_sym, nbr_s, sig, eps)
    for k in range(3):
        np.add.at(F[k], rows_sym, f[k])
        np.add.at(F[k], nbr_s, -f[k])
        F[k] = ora.ghost_put_add(F[k], n, opid)
    m = nbr != rows_full
    Ff = np.zeros((3, n))
    ff = _lj(allpos, rows_full[m], nbr[m], sig, eps)
    for k in range(3):
        np.add.at(Ff[k], rows_full[m], ff[k])
    scale = np.sqrt((Ff ** 2).sum(axis=0)).max()
    assert np.abs(F[:, :n] - Ff).max() <= 1e-12 * scale


def test_ghost_put_add_order_and_sorted_layout():
    """Values of the ghosts of one source are added in ghost order; the sorted
    layout (dst) gives the same sums at the permuted places."""
    vals = np.array([1.0, 2.0, 3.0, 1e16, -1e16, 1.0, 5.0])  # 3 domain, 4 ghosts
    opid = np.array([0, 0, 0, 2])
    out = ora.ghost_put_add(vals, 3, opid)
    assert out[0] == ((1.0 + 1e16) + -1e16) + 1.0 and out[1] == 2.0 and out[2] == 8.0
    perm = np.array([4, 0, 6, 2, 1, 5, 3])  # original -> position
    sv = np.empty_like(vals)
    sv[perm] = vals
    out2 = ora.ghost_put_add(sv, 3, opid, dst=perm)
    assert np.array_equal(out2[perm], out)


def test_sorted_plane_checks_reject_bad_tensors():
    """ghost_put_sorted and pair_lj_sym check their tensors before the kernels
    index them by sorted index up to n_part (no out-of-bounds writes)."""
    import torch

    from openfpm_pdata_amd import vector_dist as vdm

    dts = (torch.float32, torch.float64)
    with pytest.raises(ValueError, match="fewer than n_part"):
        vdm.check_sorted_planes([torch.zeros(9)], 10, dts, "put")
    with pytest.raises(ValueError, match="not contiguous"):
        vdm.check_sorted_planes([torch.zeros(40)[::2]], 10, dts, "put")
    with pytest.raises(ValueError, match="one of"):
        vdm.check_sorted_planes([torch.zeros(10, dtype=torch.int8)], 10, dts, "put")
    with pytest.raises(ValueError, match="not on the GPU"):
        vdm.check_sorted_planes([torch.zeros(10)], 10, dts, "put")
    with pytest.raises(ValueError, match="fewer than n_part"):
        vdm.check_pos_sorted(torch.zeros(3, 9), 10, torch.float32, "lj")
    with pytest.raises(ValueError, match="not contiguous"):
        vdm.check_pos_sorted(torch.zeros(20, 3).t(), 10, torch.float32, "lj")
    with pytest.raises(ValueError, match="tensor of"):
        vdm.check_pos_sorted(torch.zeros(3, 10, dtype=torch.float64), 10, torch.float32, "lj")
    with pytest.raises(ValueError, match="not on the GPU"):
        vdm.check_pos_sorted(torch.zeros(3, 10), 10, torch.float32, "lj")
