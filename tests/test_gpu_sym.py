"""Symmetric Verlet lists, the symmetric LJ pass and ghost_put<add_> on the GPU
through the C ABI: the half list bit-exact against the oracle's rule applied to
the GPU's own full list (sorted space), and the property of the reference's
vector_dist_symmetric_cell_list (Vector/tests/vector_dist_cell_list_tests.cpp:
443-649): after ghost_put<add_>, the symmetric pass's neighbour counts equal
the full pass's (exact) and its forces the full FP64 forces (1e-12 of the
largest force: the sums run in another order, through atomics)."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import _lib as L
from openfpm_pdata_amd import vector_dist as vdm
from openfpm_pdata_amd import workloads
from test_gpu_parity import make_vd

pytestmark = pytest.mark.gpu


def snap(wl):
    """Positions on multiples of 2^-20 in [0, 1): x +- 1 is exact in FP32 and FP64,
    so a ghost pair's distance is the same from both ends (no cut-off ties)."""
    q = np.round(wl.pos.astype(np.float64) * 2.0 ** 20) / 2.0 ** 20
    wl.pos[...] = np.where(q >= 1.0, q - 1.0, q).astype(wl.pos.dtype)
    return wl


def _build(dt, n, sub, seed=1):
    wl = snap(workloads.uniform_lj(n=n, dtype=dt, seed=seed))
    r = wl.r_cut * 1.1
    vd = make_vd(wl, dt, ghost=r)
    vd.ghost_get()
    cl, full, sym = vd.getVerletSymGPU(r, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION, sub=sub)
    return wl, vd, cl, full, sym


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("sub", [1, 2])
def test_sym_list_matches_oracle_rule(dt, sub):
    print(f"sym list {np.dtype(dt).name} sub {sub}", flush=True)
    wl, vd, cl, full, sym = _build(dt, 65536, sub)
    off, nbr = full.arrays()
    offs, nbrs = sym.arrays()
    a = cl.arrays()
    pid, sid = vd.ghost_opart()
    torch.cuda.synchronize()
    dom = a["dom"].cpu().numpy().astype(np.int64)
    n2s = a["n2s"].cpu().numpy().astype(np.int64)
    s2n = a["s2n"].cpu().numpy().astype(np.int64)
    gm = wl.n
    want_off, want_nbr = ora.verlet_sym(off.cpu().numpy(), nbr.cpu().numpy().astype(np.uint32), dom, n2s, gm,
                                        pid.cpu().numpy().astype(np.int64), sid.cpu().numpy().astype(np.int64),
                                        s2n=s2n)
    assert np.array_equal(offs.cpu().numpy(), want_off)
    assert np.array_equal(nbrs.cpu().numpy().astype(np.int64) & 0xFFFFFFFF, want_nbr)
    assert 2 * len(want_nbr) == int(off[-1].item()) - wl.n  # half of the non-self entries


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_sym_lj_and_ghost_put_equal_full_pass(dt):
    print(f"sym lj {np.dtype(dt).name}", flush=True)
    wl, vd, cl, full, sym = _build(dt, 65536, 2, seed=4)
    sig, eps = wl.params["sigma"], wl.params["eps"]
    F, nn = vdm.pair_lj_sym(cl, sym, vd.pos, wl.r_cut, sig, eps)
    vd.ghost_put_sorted(cl, [F[0], F[1], F[2], nn])
    # the full pass over the full list
    Ff = torch.zeros((3, vd.cap), dtype=torch.float64, device="cuda")
    U = torch.zeros(vd.cap, dtype=torch.float64, device="cuda")
    nf = torch.zeros(vd.cap, dtype=torch.int32, device="cuda")
    prm = L.LJParams(wl.r_cut, sig, eps)
    L.call("ofp_pair_lj", cl.h, full.h, vd.pos.data_ptr(), vd.cap, L.OFP_PREC_F64, prm, Ff[0].data_ptr(),
           Ff[1].data_ptr(), Ff[2].data_ptr(), U.data_ptr(), nf.data_ptr(), vdm._stream())
    torch.cuda.synchronize()
    dom = cl.arrays()["dom"].cpu().numpy().astype(np.int64)
    nn_s = nn.cpu().numpy()[dom]
    nn_f = nf.cpu().numpy()[dom]
    assert np.array_equal(nn_s, nn_f)  # both passes skip the particle itself
    Fs = F.cpu().numpy()[:, dom]
    Fr = Ff.cpu().numpy()[:, dom]
    scale = np.sqrt((Fr ** 2).sum(axis=0)).max()
    assert np.abs(Fs - Fr).max() <= 1e-12 * scale
    # a second symmetric pass on the same lists gives the same counts (outputs cleared first)
    F2, nn2 = vdm.pair_lj_sym(cl, sym, vd.pos, wl.r_cut, sig, eps)
    vd.ghost_put_sorted(cl, [nn2])
    torch.cuda.synchronize()
    assert np.array_equal(nn2.cpu().numpy()[dom], nn_s)


def test_list_kinds_are_not_mixed():
    wl, vd, cl, full, sym = _build(np.float64, 4096, 1)
    prm = L.LJParams(wl.r_cut, wl.params["sigma"], wl.params["eps"])
    F = torch.zeros((3, vd.cap), dtype=torch.float64, device="cuda")
    with pytest.raises(L.OfpError):
        L.call("ofp_pair_lj", cl.h, sym.h, vd.pos.data_ptr(), vd.cap, L.OFP_PREC_F64, prm, F[0].data_ptr(),
               F[1].data_ptr(), F[2].data_ptr(), None, None, vdm._stream())
    with pytest.raises(L.OfpError):
        vdm.pair_lj_sym(cl, full, vd.pos, wl.r_cut, 1.0, 1.0)
