"""Sparse cell list (CELLLIST_GPU_SPARSE) on the GPU through the C ABI: its
non-empty cells, starts and 27-neighbour table bit-exact against the oracle's
restatement over the oracle's dense cell list, and every particle's neighbour
candidates through the sparse table equal to the dense 27-cell walk's (the
property vector_dist_gpu_test_sparse checks, Vector/cuda/
vector_dist_gpu_unit_tests.cu:570-597)."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import vector_dist as vdm
from openfpm_pdata_amd import workloads
from test_gpu_parity import build_all, make_vd

pytestmark = pytest.mark.gpu

CASES = [("uniform", np.float64, lambda: workloads.uniform_lj(n=20000, dtype=np.float64)),
         ("clustered", np.float32, lambda: workloads.clustered_sph(n=65536))]


@pytest.mark.parametrize("name,dt,mk", CASES, ids=[c[0] for c in CASES])
def test_sparse_cell_list_matches_oracle_and_dense_walk(name, dt, mk):
    print(f"sparse cell list {name}", flush=True)
    wl = mk()
    vd, cl, allpos = build_all(wl, dt)
    sp = cl.sparse()
    dense = cl.arrays()["cell_start"].cpu().numpy().astype(np.int64)
    torch.cuda.synchronize()
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, dt)
    cs, _ = ora.cl_fill(allpos, geom)
    key, start, nn = ora.cl_sparse(cs, geom.divc)
    assert np.array_equal(sp["key"].cpu().numpy().astype(np.int64), key)
    assert np.array_equal(sp["start"].cpu().numpy().astype(np.int64), start)
    assert np.array_equal(sp["nn"].cpu().numpy().astype(np.int64), nn)
    assert len(key) < len(cs) - 1  # the padding cells hold no particle everywhere
    # neighbour candidates of the particles of a sample of non-empty cells:
    # sparse table == dense 27-cell walk (as sets and in walk order)
    d0, d1, d2 = (int(x) for x in geom.divc)
    rng = np.random.default_rng(0)
    for m in rng.choice(len(key), size=min(300, len(key)), replace=False):
        c = int(key[m])
        cx, cy, cz = c % d0, (c // d0) % d1, c // (d0 * d1)
        dense_c = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    x, y, z = cx + dx, cy + dy, cz + dz
                    if 0 <= x < d0 and 0 <= y < d1 and 0 <= z < d2:
                        q = x + d0 * (y + d1 * z)
                        dense_c.extend(range(dense[q], dense[q + 1]))
        sparse_c = []
        for e in nn[m]:
            if e >= 0:
                sparse_c.extend(range(start[e], start[e + 1]))
        assert sparse_c == dense_c


@pytest.mark.parametrize("box_nn", [2, 3])
def test_sparse_cell_list_box_nn(box_nn):
    """Cells of r_c / B with the (2B + 1)^3 stencil (setBoxNN): the sparse table
    has that stencil, bit-exact with the oracle, and its candidates equal the
    dense list's stencil walk."""
    wl = workloads.clustered_sph(n=30000, seed=17)
    dt = np.float32
    vd = make_vd(wl, dt)
    vd.ghost_get()
    allpos = np.ascontiguousarray(vd.getPosVector().cpu().numpy())
    cl = vd.getCellListGPU(wl.r_cut / box_nn, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION,
                           NNIteratorBox=box_nn)
    vd.updateCellListGPU(cl)
    sp = cl.sparse()
    assert sp["box_nn"] == box_nn
    dense = cl.arrays()["cell_start"].cpu().numpy().astype(np.int64)
    torch.cuda.synchronize()
    key, start, nn = ora.cl_sparse(dense, cl.view().divc[:], box_nn)
    assert np.array_equal(sp["key"].cpu().numpy().astype(np.int64), key)
    assert np.array_equal(sp["start"].cpu().numpy().astype(np.int64), start)
    assert np.array_equal(sp["nn"].cpu().numpy().astype(np.int64), nn)
    d0, d1, d2 = (int(x) for x in cl.view().divc)
    B = range(-box_nn, box_nn + 1)
    rng = np.random.default_rng(1)
    for m in rng.choice(len(key), size=min(100, len(key)), replace=False):
        c = int(key[m])
        cx, cy, cz = c % d0, (c // d0) % d1, c // (d0 * d1)
        dense_c = [i for dz in B for dy in B for dx in B
                   if 0 <= cx + dx < d0 and 0 <= cy + dy < d1 and 0 <= cz + dz < d2
                   for i in range(dense[cx + dx + d0 * (cy + dy + d1 * (cz + dz))],
                                  dense[cx + dx + d0 * (cy + dy + d1 * (cz + dz)) + 1])]
        assert [i for e in nn[m] if e >= 0 for i in range(start[e], start[e + 1])] == dense_c


def test_sparse_view_follows_each_construct():
    wl = workloads.uniform_lj(n=5000, dtype=np.float64)
    vd, cl, allpos = build_all(wl, np.float64)
    n1 = len(cl.sparse()["key"])
    # a second construct on half the particles gives its own sparse form
    vd2, cl2, allpos2 = build_all(workloads.uniform_lj(n=2500, dtype=np.float64, seed=9), np.float64)
    n2 = len(cl2.sparse()["key"])
    torch.cuda.synchronize()
    assert 0 < n2 <= n1
