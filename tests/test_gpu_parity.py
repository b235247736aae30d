"""GPU parity of the HIP path against the CPU oracle (bit-exact for ghosts,
cell lists and Verlet lists; FP64 physics to 1e-12 relative)."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import _lib as L
from openfpm_pdata_amd import vector_dist as vdm
from openfpm_pdata_amd import workloads

pytestmark = pytest.mark.gpu

NP2T = {np.float32: torch.float32, np.float64: torch.float64}


def make_vd(wl, dt, props=None, capacity=None, ghost=None):
    vd = vdm.VectorDistGPU(wl.n, wl.box_lo, wl.box_hi, wl.bc, ghost or wl.ghost, dtype=NP2T[dt], props=props,
                           capacity=capacity)
    vd.pos[:, : wl.n] = torch.from_numpy(np.ascontiguousarray(wl.pos.astype(dt))).cuda()
    return vd


def oracle_ghosts(wl, dt):
    g = wl.ghost
    blo, bhi, sv = ora.local_ghost_boxes(wl.box_lo, wl.box_hi, [-g] * 3, [g] * 3, wl.bc, dt)
    shifts = ora.shift_vectors(wl.box_lo, wl.box_hi, dt)
    return ora.ghost_emit(np.ascontiguousarray(wl.pos.astype(dt)), wl.n, blo, bhi, sv, shifts)


CASES = [
    ("uniform_lj", np.float64, lambda: workloads.uniform_lj(n=65536)),
    ("uniform_lj32", np.float32, lambda: workloads.uniform_lj(n=30000, dtype=np.float32, seed=99)),
    ("clustered_sph", np.float32, lambda: workloads.clustered_sph(n=50000, seed=5)),
]


# ---------------------------------------------------------------- ghosts
@pytest.mark.parametrize("name,dt,mk", CASES, ids=[c[0] for c in CASES])
def test_ghost_get_bit_exact(name, dt, mk):
    wl = mk()
    props = {"a": (torch.float32, 1), "v": (torch.float64, 3)}
    vd = make_vd(wl, dt, props)
    a = torch.arange(wl.n, dtype=torch.float32, device="cuda")
    vd.prp["a"][0, : wl.n] = a
    for k in range(3):
        vd.prp["v"][k, : wl.n] = a.double() * 10 + k
    G = vd.ghost_get(props=("a", "v"))
    gp, op, os_ = oracle_ghosts(wl, dt)
    assert G == len(op) and G > 0
    pos = vd.getPosVector().cpu().numpy()
    np.testing.assert_array_equal(pos[:, : wl.n], wl.pos.astype(dt))
    np.testing.assert_array_equal(pos[:, wl.n:], gp)
    pid, sid = vd.ghost_opart()
    np.testing.assert_array_equal(pid.cpu().numpy(), op)
    np.testing.assert_array_equal(sid.cpu().numpy(), os_)
    np.testing.assert_array_equal(vd.getProp("a")[0, wl.n:].cpu().numpy(), op.astype(np.float32))
    np.testing.assert_array_equal(vd.getProp("v")[2, wl.n:].cpu().numpy(), op * 10.0 + 2)


def test_ghost_kernels_golden_gpu():
    """The reference's kernel test boxes (vector_dist_cuda_func_test.cu:19-389)
    through ofp_ghost_create/label/emit."""
    import json
    import os

    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ghost_kat.json")))
    pos = np.asarray(g["pos"], np.float32)
    n = pos.shape[1]
    shifts = np.zeros((27, 3), np.float32)
    shifts[:4] = np.asarray(g["shifts"], np.float32)
    blo = np.asarray(g["box_lo"], np.float32)
    bhi = np.asarray(g["box_hi"], np.float32)
    sv = np.asarray(g["box_sv"], np.uint32)
    import ctypes as C

    h = C.c_void_p()
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    L.call("ofp_ghost_create", C.byref(h), L.OFP_F32, 4, p(blo), p(bhi), p(sv), p(shifts))
    try:
        tot = len(g["expected_opart"][0])
        cap = n + tot
        dev = torch.zeros((3, cap), dtype=torch.float32, device="cuda")
        dev[:, :n] = torch.from_numpy(pos).cuda()
        ng = C.c_size_t()
        L.call("ofp_ghost_label", h, dev.data_ptr(), cap, n, C.byref(ng), None)
        assert ng.value == tot
        L.call("ofp_ghost_emit", h, dev.data_ptr(), cap, n, n, None, 0, None)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dev[:, n:].cpu().numpy(), np.asarray(g["expected_ghost_pos"], np.float32))
        pid, sid, nn = C.c_void_p(), C.c_void_p(), C.c_size_t()
        L.call("ofp_ghost_opart", h, C.byref(pid), C.byref(sid), C.byref(nn))
        np.testing.assert_array_equal(vdm.device_view(pid.value, nn.value, torch.int32).cpu().numpy(),
                                      g["expected_opart"][0])
        np.testing.assert_array_equal(vdm.device_view(sid.value, nn.value, torch.int32).cpu().numpy(),
                                      g["expected_opart"][1])
    finally:
        L.lib().ofp_ghost_destroy(h)


def test_ghost_refresh_skip_labelling():
    wl = workloads.uniform_lj(n=20000, seed=21)
    vd = make_vd(wl, np.float64, {"a": (torch.float64, 1)})
    vd.prp["a"][0, : wl.n] = torch.arange(wl.n, dtype=torch.float64, device="cuda")
    G = vd.ghost_get(props=("a",))
    _, op, os_ = oracle_ghosts(wl, np.float64)
    # move particles slightly and change the property, then refresh from the cache
    vd.pos[:, : wl.n] += 1e-4
    vd.prp["a"][0, : wl.n] *= 3
    vd.ghost_get(props=("a",), opt=vdm.SKIP_LABELLING | vdm.RUN_ON_DEVICE)
    moved = wl.pos + 1e-4
    shifts = ora.shift_vectors(wl.box_lo, wl.box_hi)
    expect = moved[:, op] - shifts[os_].T
    pos = vd.getPosVector().cpu().numpy()
    np.testing.assert_array_equal(pos[:, wl.n:wl.n + G], expect)
    np.testing.assert_array_equal(vd.getProp("a")[0, wl.n:].cpu().numpy(), op * 3.0)
    # NO_POSITION: properties only
    vd.prp["a"][0, : wl.n] += 1
    vd.ghost_get(props=("a",), opt=vdm.SKIP_LABELLING | vdm.NO_POSITION)
    np.testing.assert_array_equal(vd.getProp("a")[0, wl.n:].cpu().numpy(), op * 3.0 + 1)
    np.testing.assert_array_equal(vd.getPosVector()[:, wl.n:].cpu().numpy(), expect)
    # KEEP_PROPERTIES is the same flag (util/common_pdata.hpp:12-13): a refresh
    # from the cache, positions included, with the ghost layout unchanged
    vd.pos[:, : wl.n] += 1e-4
    vd.prp["a"][0, : wl.n] += 1
    assert vd.ghost_get(props=("a",), opt=vdm.KEEP_PROPERTIES) == G
    np.testing.assert_array_equal(vd.getPosVector()[:, wl.n:wl.n + G].cpu().numpy(),
                                  (moved + 1e-4)[:, op] - shifts[os_].T)
    np.testing.assert_array_equal(vd.getProp("a")[0, wl.n:].cpu().numpy(), op * 3.0 + 2)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_map_apply_bc(dt):
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.3, 1.3, (3, 10000)).astype(dt)
    wl = workloads.Workload("bc", np.ascontiguousarray(pos), (0.0,) * 3, (1.0,) * 3, (1, 0, 1), 0.05, 0.05)
    vd = make_vd(wl, dt)
    vd.map()
    expect = ora.apply_bc(np.ascontiguousarray(pos), wl.box_lo, wl.box_hi, wl.bc)
    np.testing.assert_array_equal(vd.getPosVector().cpu().numpy(), expect)


def test_ghost_empty_and_nonperiodic():
    wl = workloads.water_column(side=12)
    vd = make_vd(wl, np.float64)
    assert vd.ghost_get() == 0
    empty = workloads.Workload("empty", np.zeros((3, 0)), (0.0,) * 3, (1.0,) * 3, (1, 1, 1), 0.1, 0.1)
    vd0 = make_vd(empty, np.float64)
    assert vd0.ghost_get() == 0


# ------------------------------------------------------------- cell list
def build_all(wl, dt, opt=vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION, props=None):
    vd = make_vd(wl, dt, props)
    vd.ghost_get(props=tuple(props or ()))
    allpos = np.ascontiguousarray(vd.getPosVector().cpu().numpy())
    cl = vd.getCellListGPU(wl.r_cut, opt)
    vd.updateCellListGPU(cl, props=tuple(props or ()))
    torch.cuda.synchronize()
    return vd, cl, allpos


@pytest.mark.parametrize("name,dt,mk", CASES, ids=[c[0] for c in CASES])
def test_cell_list_bit_exact(name, dt, mk):
    wl = mk()
    vd, cl, allpos = build_all(wl, dt, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION | vdm.CL_QUANT)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, dt)
    np.testing.assert_array_equal(cl.box_lo, geom.lo)
    np.testing.assert_array_equal(cl.div, geom.div)
    cs, ci = ora.cl_fill(allpos, geom)
    arr = cl.arrays()
    np.testing.assert_array_equal(arr["cell_start"].cpu().numpy(), cs)
    s2n = arr["s2n"].cpu().numpy()
    np.testing.assert_array_equal(s2n, ci)
    n2s = arr["n2s"].cpu().numpy()
    np.testing.assert_array_equal(n2s[s2n], np.arange(len(s2n)))
    dom = arr["dom"].cpu().numpy()
    np.testing.assert_array_equal(dom, np.nonzero(s2n < wl.n)[0])
    # reordered positions
    np.testing.assert_array_equal(vd.getPosVector().cpu().numpy(), allpos[:, s2n])
    # quantised cell-relative coordinates reproduce positions to one quantum
    q = arr["quant"].cpu().numpy().astype(np.int64)
    cid = ora.cell_ids(allpos, geom)[s2n].astype(np.int64)
    cx = cid % geom.divc[0]
    cy = (cid // geom.divc[0]) % geom.divc[1]
    cz = cid // (geom.divc[0] * geom.divc[1])
    np.testing.assert_array_equal(q[:, 3], cx | (cy << 11) | (cz << 22))
    for k, c in enumerate((cx, cy, cz)):
        sp = float(geom.sp[k])
        # stored g = u + c 2^28 (mod 2^32)
        u = ((q[:, k] - (c << 28) + 2 ** 31) % 2 ** 32) - 2 ** 31
        rec = float(geom.lo[k]) + (c - 1) * sp + u * (sp / 2 ** 28)
        assert np.abs(rec - allpos[k, s2n]).max() <= 1.01 * sp / 2 ** 29 + 1e-15


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_cell_key_half_open_faces_gpu(dt):
    """Particles on exactly representable cell faces and one ulp below them
    (the oracle's face test): the GPU bins them into the cells the half-open
    rule fixes, bit-exact with the oracle."""
    from tests.test_oracle import _Grid, face_points

    lo, sp, nk, pos, cells = face_points(dt)
    div = nk + 1
    hi = lo + div * sp
    cl = vdm.CellListGPU(NP2T[dt], tuple(lo), tuple(hi), tuple(int(d) for d in div), vdm.CL_NON_SYMMETRIC,
                         float(sp.min()))
    n = pos.shape[1]
    t = torch.from_numpy(pos).cuda()
    cl.construct(t, n, n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(cl.view().spacing[:], sp)
    divc = div + 2
    arr = cl.arrays()
    cs = arr["cell_start"].cpu().numpy().astype(np.int64)
    s2n = arr["s2n"].cpu().numpy().astype(np.int64)
    cid = np.empty(n, np.int64)
    cid[s2n] = np.searchsorted(cs, np.arange(n), side="right") - 1
    got = np.stack([cid % divc[0], (cid // divc[0]) % divc[1], cid // (divc[0] * divc[1])], axis=1)
    np.testing.assert_array_equal(got, cells)
    np.testing.assert_array_equal(cid, ora.cell_ids(pos, _Grid(lo, sp, divc, dt)))


def test_cell_list_dense_cells():
    """Cells above the small-cell limit (64) and above one LDS chunk (4096)
    keep the canonical ascending order."""
    rng = np.random.default_rng(8)
    n = 30000
    pos = rng.random((3, n))
    pos[:, :9000] = 0.5 + rng.random((3, 9000)) * 0.004  # one cell with > 4096 particles
    pos[:, 9000:12000] = 0.2 + rng.random((3, 3000)) * 0.01  # dense but < 4096
    rng.shuffle(pos, axis=1)
    wl = workloads.Workload("dense", np.ascontiguousarray(pos.astype(np.float32)), (0.0,) * 3, (1.0,) * 3,
                            (1, 1, 1), 0.05, 0.05)
    vd, cl, allpos = build_all(wl, np.float32)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, np.float32)
    cs, ci = ora.cl_fill(allpos, geom)
    assert np.diff(cs.astype(np.int64)).max() > 4096
    arr = cl.arrays()
    np.testing.assert_array_equal(arr["s2n"].cpu().numpy(), ci)
    np.testing.assert_array_equal(arr["cell_start"].cpu().numpy(), cs)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("swaps", [0.0, 0.02])
@pytest.mark.parametrize("bc", [(1, 1, 1), (0, 0, 0)])
def test_cell_list_presorted_input(dt, swaps, bc):
    """Input already in cell order (a run that keeps its particles sorted), exact
    or with a few random transpositions: the build that gathers the positions
    once (instead of carrying them through the passes) gives the same result.
    Non-periodic: no ghost tail, so the input stays near cell order."""
    wl = workloads.clustered_sph(n=60000, seed=12)
    wl.bc = bc
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, dt)
    pos = wl.pos.astype(dt)
    order = np.argsort(ora.cell_ids(np.ascontiguousarray(pos), geom), kind="stable")
    pos = pos[:, order]
    rng = np.random.default_rng(13)
    k = int(swaps * wl.n)
    if k:
        a, b = rng.integers(0, wl.n, k), rng.integers(0, wl.n, k)
        pos[:, a], pos[:, b] = pos[:, b].copy(), pos[:, a].copy()
    wl2 = workloads.Workload("presorted", np.ascontiguousarray(pos), wl.box_lo, wl.box_hi, wl.bc, wl.ghost, wl.r_cut)
    vd, cl, allpos = build_all(wl2, dt, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION | vdm.CL_QUANT)
    cs, ci = ora.cl_fill(allpos, geom)
    arr = cl.arrays()
    np.testing.assert_array_equal(arr["cell_start"].cpu().numpy(), cs)
    s2n = arr["s2n"].cpu().numpy()
    np.testing.assert_array_equal(s2n, ci)
    np.testing.assert_array_equal(arr["n2s"].cpu().numpy()[s2n], np.arange(len(s2n)))
    np.testing.assert_array_equal(arr["dom"].cpu().numpy(), np.nonzero(s2n < wl.n)[0])
    np.testing.assert_array_equal(vd.getPosVector().cpu().numpy(), allpos[:, s2n])


def test_reorder_restore_identity():
    """CL_GPU_REORDER + restoreOrder is the identity (vector_dist_gpu_unit_tests.cu:425-568)."""
    wl = workloads.clustered_sph(n=40000, seed=6)
    props = {"a": (torch.float32, 1), "v": (torch.float64, 3)}
    vd = make_vd(wl, np.float32, props)
    vd.prp["a"][0, : wl.n] = torch.randn(wl.n, device="cuda")
    vd.prp["v"][:, : wl.n] = torch.randn(3, wl.n, device="cuda", dtype=torch.float64)
    vd.ghost_get(props=("a", "v"))
    before = {k: vd.getProp(k).clone() for k in ("a", "v")}
    pos_before = vd.getPosVector().clone()
    cl = vd.getCellListGPU(wl.r_cut, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER)
    vd.updateCellListGPU(cl, props=("a", "v"))
    s2n = cl.arrays()["s2n"].long()
    assert torch.equal(vd.getProp("a"), before["a"][:, s2n])
    vd.restoreOrder(cl, props=("a", "v"))
    assert torch.equal(vd.getPosVector(), pos_before)
    for k in before:
        assert torch.equal(vd.getProp(k), before[k])


# ---------------------------------------------------------------- Verlet
def _verlet_case(wl, dt, r, sub=1):
    vd = make_vd(wl, dt)
    vd.ghost_get()
    allpos = np.ascontiguousarray(vd.getPosVector().cpu().numpy())
    cl, vl = vd.getVerletGPU(r, sub=sub)
    torch.cuda.synchronize()
    off, nbr = vl.arrays()
    arr = cl.arrays()
    s2n = arr["s2n"].cpu().numpy().astype(np.int64)
    dom = arr["dom"].cpu().numpy().astype(np.int64)
    return vd, allpos, off.cpu().numpy(), nbr.cpu().numpy().astype(np.int64), s2n, dom


@pytest.mark.parametrize("name,dt,mk", CASES, ids=[c[0] for c in CASES])
def test_verlet_bit_exact(name, dt, mk):
    wl = mk()
    r = wl.r_cut  # cell = r; the oracle walks the same geometry as getVerlet
    vd, allpos, off, nbr, s2n, dom = _verlet_case(wl, dt, r)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, r, vdm.GHOST_ENLARGE_VL, dt)
    ooff, onbr = ora.verlet(allpos, wl.n, geom, r)
    assert off[-1] == ooff[-1]
    assert_rows_equal_oracle(off, nbr, s2n, dom, ooff, onbr)


def assert_rows_equal_oracle(off, nbr, s2n, dom, ooff, onbr):
    """Every row: row k of the GPU list is particle s2n[dom[k]]; its entries,
    mapped through s2n, equal the oracle's row of that particle in the reference
    traversal order."""
    off = np.asarray(off, np.int64)
    ooff = np.asarray(ooff, np.int64)
    order = np.argsort(s2n[dom])  # GPU row of original particle i
    lens = np.diff(off)[order]
    np.testing.assert_array_equal(lens, np.diff(ooff))
    first = off[:-1][order]
    idx = np.repeat(first - (np.cumsum(lens) - lens), lens) + np.arange(int(lens.sum()))
    np.testing.assert_array_equal(s2n[nbr[idx]], np.asarray(onbr, np.int64))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_verlet_dilute_rows_kernel(dt):
    """A dilute list (fewer than a wave of candidates per target on average):
    the thread-per-row kernels build it, and every row equals the oracle's."""
    wl = workloads.uniform_lj(n=3000, seed=31, dtype=dt)
    r = wl.r_cut
    vd, allpos, off, nbr, s2n, dom = _verlet_case(wl, dt, r)
    n_cells = np.prod(vd.getCellListGPU(r, ghostEnlargeFactor=vdm.GHOST_ENLARGE_VL).div.astype(np.int64) + 2)
    assert vd.size * 27 < 64 * n_cells  # the per-row path
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, r, vdm.GHOST_ENLARGE_VL, dt)
    ooff, onbr = ora.verlet(allpos, wl.n, geom, r)
    assert_rows_equal_oracle(off, nbr, s2n, dom, ooff, onbr)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("sub", [1, 2, 3])
def test_verlet_with_skin_vs_brute_force(dt, sub):
    """Rows as sets (ascending original index) equal the O(N^2) neighbour sets
    (vector_dist_NN_tests.cpp:213-418) for cells of r / sub and a sub-cell
    stencil."""
    wl = workloads.clustered_sph(n=6000, seed=12, dtype=dt)
    r = wl.params["r_verlet"]
    wl.ghost = r
    vd, allpos, off, nbr, s2n, dom = _verlet_case(wl, dt, r, sub)
    boff, bnbr = ora.brute_nn(allpos, wl.n, r)
    orig_of_row = s2n[dom]
    for k in range(len(dom)):
        i = orig_of_row[k]
        np.testing.assert_array_equal(np.sort(s2n[nbr[off[k]:off[k + 1]]]), bnbr[boff[i]:boff[i + 1]])


def test_verlet_lattice_shells():
    """vector_dist_unit_test.cpp:1289-1318: 7 / 12 / 8."""
    import math

    ng = 24
    dt = np.float32
    pos = workloads.lattice(ng, dt)
    sp = dt(1.0) / dt(ng)
    first, second, third = sp, dt(math.sqrt(2.0 * sp * sp)), dt(math.sqrt(3.0 * sp * sp))
    wl = workloads.Workload("lattice", pos, (0.0,) * 3, (1.0,) * 3, (1, 1, 1), float(third * dt(1.1)), 0.0)
    r = float(dt(third + third * dt(0.05)))
    vd, allpos, off, nbr, s2n, dom = _verlet_case(wl, dt, r)
    sorted_pos = vd.getPosVector().cpu().numpy()
    f1 = dt(first + first * dt(0.05))
    f2 = dt(second + second * dt(0.05))
    for k in range(len(dom)):
        row = nbr[off[k]:off[k + 1]]
        d = np.sqrt(((sorted_pos[:, row] - sorted_pos[:, [dom[k]]]) ** 2).sum(axis=0))
        assert (d <= f1).sum() == 7 and ((d > f1) & (d <= f2)).sum() == 12 and (d > f2).sum() == 8


# --------------------------------------------------------------- physics
def norm_err(got, ref):
    """Per-particle error |F - F64| / max(|F64_i|, RMS_j |F64_j|); returns (max, rms)."""
    got = np.atleast_2d(got.T).T if got.ndim == 1 else got
    ref = np.atleast_2d(ref.T).T if ref.ndim == 1 else ref
    d = np.sqrt(((got - ref) ** 2).sum(axis=1))
    mag = np.sqrt((ref ** 2).sum(axis=1))
    rms = np.sqrt((mag ** 2).mean())
    e = d / np.maximum(mag, rms)
    return float(e.max()), float(np.sqrt((e ** 2).mean()))


def run_lj(wl, dt, precision, verlet=False):
    opt = vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION | vdm.CL_QUANT
    vlh = None
    if verlet:
        # Verlet rows at 1.1 r_c (the masked row walker), ghosts to 1.1 r_c
        r = wl.r_cut * 1.1
        vd = make_vd(wl, dt, ghost=r)
        vd.ghost_get()
        allpos = np.ascontiguousarray(vd.getPosVector().cpu().numpy())
        cl, vl = vd.getVerletGPU(r, opt, sub=2)
        vlh = vl.h
    else:
        vd, cl, allpos = build_all(wl, dt, opt)
    odt = torch.float64 if precision == L.OFP_PREC_F64 else torch.float32
    F = torch.zeros((3, vd.cap), dtype=odt, device="cuda")
    U = torch.zeros(vd.cap, dtype=odt, device="cuda")
    nn = torch.zeros(vd.cap, dtype=torch.int32, device="cuda")
    prm = L.LJParams(wl.r_cut, wl.params["sigma"], wl.params["eps"])
    fp = vdm.planes_of(F)
    L.call("ofp_pair_lj", cl.h, vlh, vd.pos.data_ptr(), vd.cap, precision, prm, fp[0][0], fp[1][0], fp[2][0],
           U.data_ptr(), nn.data_ptr(), vdm._stream())
    torch.cuda.synchronize()
    n2s = cl.arrays()["n2s"].cpu().numpy().astype(np.int64)[: wl.n]
    return (F.cpu().numpy()[:, n2s].T.astype(np.float64), U.cpu().numpy()[n2s].astype(np.float64),
            nn.cpu().numpy()[n2s], allpos, vd, cl)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_lj_fp64_matches_oracle(dt):
    wl = workloads.uniform_lj(n=65536, dtype=dt)
    F, U, nn, allpos, _, _ = run_lj(wl, dt, L.OFP_PREC_F64)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, dt)
    oF, oU, onn = ora.lj(allpos, wl.n, geom, wl.r_cut, wl.params["sigma"], wl.params["eps"])
    np.testing.assert_array_equal(nn, onn)
    emax, _ = norm_err(F, oF)
    assert emax < 1e-12
    np.testing.assert_allclose(U, oU, rtol=1e-11, atol=1e-12 * np.abs(oU).max())


def test_lj_fp32c_error_bound():
    """FP32 quantised cell-relative records (pairs evaluated in FP64) vs the FP64
    oracle; tolerance: max normalised error 1e-6, RMS 1e-7."""
    wl = workloads.uniform_lj(n=65536)
    F, U, nn, allpos, _, _ = run_lj(wl, np.float64, L.OFP_PREC_F32C)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, np.float64)
    oF, oU, onn = ora.lj(allpos, wl.n, geom, wl.r_cut, wl.params["sigma"], wl.params["eps"])
    emax, erms = norm_err(F, oF)
    print(f"LJ FP32C vs FP64 oracle: max {emax:.3e} rms {erms:.3e}; nn mismatches {(nn != onn).sum()}")
    # cut-off ties are decided on the stored positions: the same pairs as FP64
    np.testing.assert_array_equal(nn, onn)
    assert emax < 1e-6 and erms < 1e-7


@pytest.mark.parametrize("precision", [L.OFP_PREC_F64, L.OFP_PREC_F32C])
def test_lj_verlet_rows_match_oracle(precision):
    """LJ over Verlet rows at 1.1 r_c (the row walker evaluates every entry and
    masks the misses and self): exact neighbour counts in both forms, FP64
    1e-12, FP32C the 1e-6 / 1e-7 bound; every value finite."""
    wl = workloads.uniform_lj(n=65536)
    F, U, nn, allpos, _, _ = run_lj(wl, np.float64, precision, verlet=True)
    assert np.isfinite(F).all() and np.isfinite(U).all()
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.r_cut * 1.1, wl.r_cut, vdm.GHOST_ENLARGE_CL, np.float64)
    oF, oU, onn = ora.lj(allpos, wl.n, geom, wl.r_cut, wl.params["sigma"], wl.params["eps"])
    emax, erms = norm_err(F, oF)
    np.testing.assert_array_equal(nn, onn)
    if precision == L.OFP_PREC_F64:
        assert emax < 1e-12
    else:
        assert emax < 1e-6 and erms < 1e-7


def run_sph(wl, dt, precision, verlet=False, sub=1):
    opt = vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION | vdm.CL_QUANT
    pdt = torch.float64 if precision == L.OFP_PREC_F64 else torch.float32
    props = {"rho": (torch.float64, 1), "pterm": (pdt, 1)}
    vd = make_vd(wl, dt, props)
    vd.ghost_get()
    allpos = np.ascontiguousarray(vd.getPosVector().cpu().numpy())
    vlh = None
    if verlet:
        cl, vl = vd.getVerletGPU(wl.r_cut, opt, sub=sub)
        vlh = vl.h
    else:
        cl = vd.getCellListGPU(wl.r_cut / sub, opt, NNIteratorBox=sub)
        vd.updateCellListGPU(cl)
    prm = L.SPHParams(wl.params["h"], wl.params["mass"], wl.params["rho0"], wl.params["c0"])
    rho = vd.prp["rho"]
    pterm = vd.prp["pterm"]
    L.call("ofp_sph_density", cl.h, vlh, vd.pos.data_ptr(), vd.cap, precision, prm, rho.data_ptr(), vdm._stream())
    # ghosts take their source's density (sorted-space ghost_get)
    vd.ghost_sync_sorted(cl, ("rho",))
    L.call("ofp_sph_eos", precision, rho.data_ptr(), vd.size, prm, None, pterm.data_ptr(), vdm._stream())
    F = torch.zeros((3, vd.cap), dtype=pdt, device="cuda")
    fp = vdm.planes_of(F)
    L.call("ofp_sph_force", cl.h, vlh, vd.pos.data_ptr(), vd.cap, precision, prm, pterm.data_ptr(),
           fp[0][0], fp[1][0], fp[2][0], vdm._stream())
    torch.cuda.synchronize()
    arr = cl.arrays()
    n2s = arr["n2s"].cpu().numpy().astype(np.int64)
    rho_np = rho[0].cpu().numpy().astype(np.float64)
    return (rho_np[n2s[: wl.n]], F.cpu().numpy()[:, n2s[: wl.n]].T.astype(np.float64), allpos,
            rho_np[n2s[: vd.size]])


def oracle_sph(wl, dt, allpos):
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, dt)
    h, m = wl.params["h"], wl.params["mass"]
    rho = ora.sph_density(allpos, wl.n, geom, h, m)
    _, op, _ = oracle_ghosts(wl, dt)
    rho_all = np.concatenate([rho, rho[op]])
    prs_all = ora.tait(rho_all, wl.params["rho0"], wl.params["c0"])
    F = ora.sph_force(allpos, wl.n, geom, h, m, rho_all, prs_all)
    return rho, F, rho_all


SPH_MODES = [(False, 1), (True, 1), (True, 2), (False, 2)]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("verlet,sub", SPH_MODES, ids=["cell", "verlet", "verlet-sub2", "cell-sub2"])
def test_sph_fp64_matches_oracle(dt, verlet, sub):
    wl = workloads.clustered_sph(n=40000, seed=17, dtype=dt)
    rho, F, allpos, rho_all_gpu = run_sph(wl, dt, L.OFP_PREC_F64, verlet, sub)
    orho, oF, orho_all = oracle_sph(wl, dt, allpos)
    np.testing.assert_allclose(rho, orho, rtol=1e-12)
    np.testing.assert_allclose(rho_all_gpu, orho_all, rtol=1e-12)
    emax, _ = norm_err(F, oF)
    assert emax < 1e-10


@pytest.mark.parametrize("verlet,sub", SPH_MODES, ids=["cell", "verlet", "verlet-sub2", "cell-sub2"])
def test_sph_fp32c_error_bound(verlet, sub):
    """Tolerance: density max relative error 1e-6; force max normalised error
    1e-6 (FP32 cell-relative records and P/rho^2, pairs evaluated in FP64)."""
    wl = workloads.clustered_sph(n=40000, seed=17)
    rho, F, allpos, _ = run_sph(wl, np.float32, L.OFP_PREC_F32C, verlet, sub)
    orho, oF, _ = oracle_sph(wl, np.float32, allpos)
    erho = np.abs(rho - orho) / np.abs(orho)
    emax, erms = norm_err(F, oF)
    print(f"SPH FP32C: rho max {erho.max():.3e} rms {np.sqrt((erho**2).mean()):.3e}; F max {emax:.3e} rms {erms:.3e}")
    assert erho.max() < 1e-6
    assert emax < 1e-6


# ------------------------------------------------------------ reductions
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_reduce(dt):
    import ctypes as C

    x = torch.randn(1_000_003, dtype=NP2T[dt], device="cuda")
    ref = x.double().cpu().numpy()
    for op, f in ((L.OFP_RED_SUM, np.sum), (L.OFP_RED_MAX, np.max), (L.OFP_RED_MIN, np.min)):
        r = C.c_double()
        L.call("ofp_reduce", {np.float32: L.OFP_F32, np.float64: L.OFP_F64}[dt], x.data_ptr(), None, x.numel(), op,
               C.byref(r), vdm._stream())
        assert abs(r.value - f(ref)) <= 1e-9 * max(1.0, np.abs(ref).sum() if op == L.OFP_RED_SUM else abs(f(ref)))


# ----------------------------------------------- sizes at the kernel edges
@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 8193])
def test_path_at_wave_and_tile_edges(n):
    """Ghosts, cell list (reorder + quantised stream), Verlet list and the FP64
    LJ pass on n uniform FP32 particles, n around the 64-lane wave and the
    4096-particle radix tile (and empty): bit-exact cells and rows, forces to
    1e-12 against the oracle."""
    dt = np.float32
    wl = workloads.uniform_lj(n=n, dtype=dt, seed=1000 + n, r_cut=0.1, sigma=0.04)
    vd, cl, allpos = build_all(wl, dt, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION | vdm.CL_QUANT)
    arr = cl.arrays()
    s2n = arr["s2n"].cpu().numpy().astype(np.int64)
    assert len(s2n) == allpos.shape[1]
    if allpos.shape[1]:
        geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, dt)
        cs, ci = ora.cl_fill(allpos, geom)
        np.testing.assert_array_equal(arr["cell_start"].cpu().numpy(), cs)
        np.testing.assert_array_equal(s2n, ci)
        np.testing.assert_array_equal(vd.getPosVector().cpu().numpy(), allpos[:, s2n])
    else:
        assert int(arr["cell_start"].cpu().numpy().max()) == 0
    vd.restoreOrder(cl)
    # Verlet rows at r_c (cells of r_c, the getVerlet geometry)
    vd2, allpos2, off, nbr, s2n2, dom = _verlet_case(wl, dt, wl.r_cut)
    assert len(off) == n + 1
    if n:
        geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_VL, dt)
        ooff, onbr = ora.verlet(allpos2, wl.n, geom, wl.r_cut)
        assert off[-1] == ooff[-1]
        rows = np.argsort(s2n2[dom])
        for i in range(n):
            k = rows[i]
            np.testing.assert_array_equal(s2n2[nbr[off[k]:off[k + 1]]], onbr[ooff[i]:ooff[i + 1]])
    else:
        assert int(off[-1]) == 0
    # LJ, FP64 arithmetic, cell walk
    F, U, nn, allpos3, _, _ = run_lj(wl, dt, L.OFP_PREC_F64)
    if n:
        geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, dt)
        oF, oU, onn = ora.lj(allpos3, wl.n, geom, wl.r_cut, wl.params["sigma"], wl.params["eps"])
        np.testing.assert_array_equal(nn, onn)
        if onn.sum():
            assert norm_err(F, oF)[0] < 1e-12
