"""map on one rank, CPU checks of the oracle: the processor of every particle
equals brute force over the sub-domains (unique for a partition, with the
domain's upper face closed in non-periodic dimensions and particles beyond a
non-periodic face lost), the wrap applied first, and the destination order on
the reference's reorder_lbl known answer (Vector/cuda/vector_dist_cuda_func_test.cu:
810-846: outputs i*10 + j of processor j land at j*10 + i)."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import remote as R

DECOMPS = [
    ((2, 2, 2), 0, (1, 1, 1)),
    ((3, 3, 3), 13, (1, 1, 1)),
    ((1, 2, 4), 3, (1, 1, 1)),
    ((2, 2, 2), 5, (0, 0, 0)),
    ((4, 1, 3), 2, (1, 0, 1)),
]


def map_points(n, seed, dt, bc, lo=0.0, hi=1.0):
    """Particles of a step that moved them: most inside, some beyond every face
    (wrapped where periodic, lost where not), some exactly on sub-domain and
    domain faces."""
    rng = np.random.default_rng(seed)
    pos = (lo - 0.05 + rng.random((3, n)) * (hi - lo + 0.1)).astype(dt)
    if n >= 40:
        for k in range(3):
            pos[k, 4 * k:4 * k + 4] = dt(hi)  # on the domain's upper face
            pos[k, 12 + 4 * k:16 + 4 * k] = dt(0.5)  # on an inner face of the 2-cuts
            pos[k, 24 + 4 * k:28 + 4 * k] = dt(lo)
    return np.ascontiguousarray(pos)


def brute_proc(w, slo, shi, sp, dom_hi, bc):
    out = np.full(w.shape[1], 0xFFFFFFFF, np.uint32)
    for p in range(w.shape[1]):
        x = w[:, p]
        hits = []
        for b in range(slo.shape[0]):
            ok = True
            for k in range(3):
                closed = not bc[k] and shi[b, k] == dom_hi[k]
                ok = ok and slo[b, k] <= x[k] and (x[k] <= shi[b, k] if closed else x[k] < shi[b, k])
            if ok:
                hits.append(int(sp[b]))
        assert len(hits) <= 1  # a partition: at most one sub-domain contains a point
        if hits:
            out[p] = hits[0]
    return out


@pytest.mark.parametrize("dims,rank,bc", DECOMPS)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_map_proc_equals_brute_force(dims, rank, bc, dt):
    tdt = torch.float32 if dt == np.float32 else torch.float64
    slo, shi, sp = R.regular_subdomains(tdt, (0, 0, 0), (1, 1, 1), dims)
    n = 3000
    pos = map_points(n, 7, dt, bc)
    dlo, dhi = np.zeros(3, dt), np.ones(3, dt)
    w, proc, order, n_stay, stay_begin, n_lost = ora.map_label(pos, n, slo, shi, sp, dlo, dhi, bc, rank)
    assert np.array_equal(w, ora.apply_bc(pos, dlo, dhi, bc))
    for k in range(3):
        if bc[k]:
            assert np.all((w[k] >= 0) & (w[k] < 1))
    assert np.array_equal(proc, brute_proc(w, slo, shi, sp, dhi, bc))
    lost = proc == 0xFFFFFFFF
    assert n_lost == int(lost.sum())
    if all(bc):
        assert n_lost == 0
    else:
        assert n_lost > 0  # particles beyond a non-periodic face
        nonper = [k for k in range(3) if not bc[k]]
        assert np.all(np.any([(w[k] < 0) | (w[k] > 1) for k in nonper], axis=0)[lost])
    # order: processor ascending, particle index inside, the lost last
    key = proc[order.astype(np.int64)]
    assert np.all(np.diff(key.astype(np.int64)) >= 0)
    for q in np.unique(proc):
        sel = order[key == q].astype(np.int64)
        assert np.all(np.diff(sel) > 0)
    assert n_stay == int(np.sum(proc == rank)) and stay_begin == int(np.sum(proc < rank))


def test_map_order_reorder_lbl_known_answer():
    """Ten slabs along x, processor j owns slab j; particle k = i*10 + j sits in
    slab j, so the destination order puts particle i*10 + j at j*10 + i."""
    f = np.float64
    slo = np.array([[j / 10, 0, 0] for j in range(10)], f)
    shi = np.array([[(j + 1) / 10, 1, 1] for j in range(10)], f)
    sp = np.arange(10, dtype=np.uint32)
    pos = np.zeros((3, 100), f)
    for i in range(10):
        for j in range(10):
            pos[:, i * 10 + j] = ((j + 0.5) / 10, 0.5, 0.5)
    w, proc, order, n_stay, stay_begin, n_lost = ora.map_label(pos, 100, slo, shi, sp, np.zeros(3), np.ones(3),
                                                             (1, 1, 1), 4)
    for i in range(10):
        for j in range(10):
            assert int(order[j * 10 + i]) == i * 10 + j
    assert (n_stay, stay_begin, n_lost) == (10, 40, 0)


def test_regular_subdomains_partition_the_box():
    slo, shi, sp = R.regular_subdomains(torch.float64, (0, 0, 0), (1, 2, 3), (2, 3, 4))
    assert slo.shape == (24, 3) and np.array_equal(sp, np.arange(24))
    vol = np.prod(shi - slo, axis=1)
    assert abs(vol.sum() - 6.0) < 1e-12
    assert np.all(shi.max(axis=0) == (1, 2, 3)) and np.all(slo.min(axis=0) == 0)
