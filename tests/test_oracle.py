"""Pins of the CPU oracle (no GPU): golden vectors from the reference's own
tests, closed forms, brute force.  Each test names the reference test or the
mathematical fact it restates."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as ora
from openfpm_pdata_amd import workloads

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name + ".json")) as f:
        return json.load(f)


def comb_lin(c):
    return (c[0] + 1) + 3 * (c[1] + 1) + 9 * (c[2] + 1)


# ----------------------------------------------------------------- geometry
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_shift_vectors_golden(dt):
    """shift_vect_converter_tests.cpp:17-51."""
    g = load("shift_vectors")
    sv = ora.shift_vectors(g["domain_lo"], g["domain_hi"], dt)
    assert sv.shape == (g["n_shifts"], 3)
    for case in g["cases"]:
        np.testing.assert_array_equal(sv[comb_lin(case["c"])], np.asarray(case["shift"], dt))
    # the zero comb is the identity shift
    np.testing.assert_array_equal(sv[comb_lin([0, 0, 0])], np.zeros(3, dt))


def test_shift_vectors_nonunit_domain():
    sv = ora.shift_vectors([-1.0, 2.0, 0.5], [3.0, 2.5, 1.5])
    L = np.array([4.0, 0.5, 1.0])
    for l in range(27):
        c = [l % 3 - 1, (l // 3) % 3 - 1, l // 9 - 1]
        np.testing.assert_array_equal(sv[l], -np.asarray(c) * L)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_local_ghost_boxes_structure(dt):
    """ie_loc_ghost: 26 boxes for a fully periodic single sub-domain: 6 faces,
    12 edges, 8 corners; box for comb c spans [lo, lo+g) where c_k = +1,
    [hi-g, hi) where c_k = -1, the full domain where c_k = 0."""
    g = 0.05
    blo, bhi, sv = ora.local_ghost_boxes([0, 0, 0], [1, 1, 1], [-g] * 3, [g] * 3, [1, 1, 1], dt)
    assert len(sv) == 26
    assert list(sv) == sorted(sv)  # createShiftBox sorts by shift id
    nz = []
    for b in range(26):
        c = [int(sv[b]) % 3 - 1, (int(sv[b]) // 3) % 3 - 1, int(sv[b]) // 9 - 1]
        nz.append(sum(1 for v in c if v != 0))
        for k in range(3):
            if c[k] == 1:
                assert blo[b, k] == 0 and bhi[b, k] == dt(g)
            elif c[k] == -1:
                assert blo[b, k] == dt(dt(1) - dt(g)) and bhi[b, k] == 1
            else:
                assert blo[b, k] == 0 and bhi[b, k] == 1
    assert sorted(nz).count(1) == 6 and nz.count(2) == 12 and nz.count(3) == 8


def test_local_ghost_boxes_partial_periodicity():
    blo, bhi, sv = ora.local_ghost_boxes([0, 0, 0], [1, 1, 1], [-0.1] * 3, [0.1] * 3, [1, 0, 1])
    assert len(sv) == 8  # combs with c_y = 0 only: 3*3 - 1
    for s in sv:
        assert (int(s) // 3) % 3 == 1
    _, _, sv0 = ora.local_ghost_boxes([0, 0, 0], [1, 1, 1], [-0.1] * 3, [0.1] * 3, [0, 0, 0])
    assert len(sv0) == 0


def test_ghost_kernels_golden():
    """vector_dist_cuda_func_test.cu:19-389 known answers (counts, ghost
    positions, o_part_loc order)."""
    g = load("ghost_kat")
    pos = np.ascontiguousarray(np.asarray(g["pos"], np.float32))
    n = pos.shape[1]
    shifts = np.zeros((27, 3), np.float32)
    shifts[:4] = np.asarray(g["shifts"], np.float32)
    cnt, tot = ora.ghost_count(pos, n, g["box_lo"], g["box_hi"], g["box_sv"])
    np.testing.assert_array_equal(cnt, g["expected_count"])
    gp, op, os_ = ora.ghost_emit(pos, n, g["box_lo"], g["box_hi"], g["box_sv"], shifts)
    assert tot == len(g["expected_opart"][0])
    np.testing.assert_array_equal(op, g["expected_opart"][0])
    np.testing.assert_array_equal(os_, g["expected_opart"][1])
    np.testing.assert_array_equal(gp, np.asarray(g["expected_ghost_pos"], np.float32))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_ghost_set_invariants(dt):
    """vector_dist_unit_test.cpp:631-729: every ghost lies outside the domain
    and inside the domain enlarged by the ghost; multiplicity 1/3/7 for
    face/edge/corner particles; E[G] = N((1+2g)^3 - 1)."""
    wl = workloads.uniform_lj(n=20000, seed=7, dtype=dt)
    g = 0.05
    blo, bhi, sv = ora.local_ghost_boxes(wl.box_lo, wl.box_hi, [-g] * 3, [g] * 3, wl.bc, dt)
    shifts = ora.shift_vectors(wl.box_lo, wl.box_hi, dt)
    gp, op, os_ = ora.ghost_emit(wl.pos, wl.n, blo, bhi, sv, shifts)
    inside = np.all((gp >= 0) & (gp < 1), axis=0)
    assert not inside.any()
    assert np.all(gp >= -dt(g) - 1e-7) and np.all(gp < 1 + dt(g) + 1e-7)
    # multiplicity from the number of near faces
    near = ((wl.pos < dt(g)) | (wl.pos >= dt(dt(1) - dt(g)))).sum(axis=0)
    mult = np.bincount(op, minlength=wl.n)
    np.testing.assert_array_equal(mult, (2 ** near) - 1)
    expect = wl.n * ((1 + 2 * g) ** 3 - 1)
    assert abs(len(op) - expect) < 5 * math.sqrt(expect)
    # ghost = source - shift exactly
    np.testing.assert_array_equal(gp, (wl.pos[:, op] - shifts[os_].T).astype(dt))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_cell_param_invariants(dt):
    """cl_param_calculate reconstruction: cells at least r_cut wide, the box
    covers the domain plus the magnified ghost, one padding layer."""
    for rc in (0.05, 0.0161, 0.031, 0.079):
        geom = ora.CellGeom((0, 0, 0), (1, 1, 1), rc, rc, float(np.float32(1.013)), dt)
        assert np.all(geom.sp >= dt(rc) * (1 - 4 * np.finfo(dt).eps))
        assert np.all(geom.lo <= -dt(rc)) and np.all(geom.hi >= 1 + dt(rc))
        assert np.all(geom.div == np.floor((1 + 2 * 1.013 * rc) / rc + 1e-9).astype(int) + 1)


def test_cell_ids_domain_particles_in_interior_cells():
    wl = workloads.uniform_lj(n=5000, seed=3)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, float(np.float32(1.013)), np.float64)
    cid = ora.cell_ids(wl.pos, geom)
    cx = cid % geom.divc[0]
    cy = (cid // geom.divc[0]) % geom.divc[1]
    cz = cid // (geom.divc[0] * geom.divc[1])
    for c, k in ((cx, 0), (cy, 1), (cz, 2)):
        assert c.min() >= 1 and c.max() <= geom.div[k]
        # floor((x - lo)/sp) + 1 closed form
        np.testing.assert_array_equal(c, np.floor((wl.pos[k] - geom.lo[k]) / geom.sp[k]).astype(int) + 1)


def test_cl_fill_canonical_order():
    wl = workloads.clustered_sph(n=6000, seed=9)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, float(np.float32(1.013)), np.float32)
    cs, ci = ora.cl_fill(wl.pos, geom)
    cid = ora.cell_ids(wl.pos, geom)
    np.testing.assert_array_equal(np.sort(ci), np.arange(wl.n))
    # stable sort by cell id == ascending index inside every cell
    np.testing.assert_array_equal(ci, np.argsort(cid, kind="stable"))
    np.testing.assert_array_equal(np.diff(cs.astype(np.int64)), np.bincount(cid, minlength=geom.n_cells))


# ------------------------------------------------------------- neighbours
def _with_ghosts(wl, dt):
    g = wl.ghost
    blo, bhi, sv = ora.local_ghost_boxes(wl.box_lo, wl.box_hi, [-g] * 3, [g] * 3, wl.bc, dt)
    shifts = ora.shift_vectors(wl.box_lo, wl.box_hi, dt)
    gp, op, os_ = ora.ghost_emit(wl.pos, wl.n, blo, bhi, sv, shifts)
    return np.ascontiguousarray(np.concatenate([wl.pos, gp], axis=1)), op


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_verlet_equals_brute_force(dt):
    """vector_dist_NN_tests.cpp:213-418: cell-list Verlet == O(N^2), self included."""
    wl = workloads.uniform_lj(n=3000, seed=11, dtype=dt, r_cut=0.08)
    allp, _ = _with_ghosts(wl, dt)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, 1.013, dt)
    off, nbr = ora.verlet(allp, wl.n, geom, wl.r_cut)
    boff, bnbr = ora.brute_nn(allp, wl.n, wl.r_cut)
    np.testing.assert_array_equal(off, boff)
    for p in range(wl.n):
        row = nbr[off[p]:off[p + 1]]
        assert p in row
        np.testing.assert_array_equal(np.sort(row), bnbr[boff[p]:boff[p + 1]])


def test_lattice_shells_golden():
    """vector_dist_unit_test.cpp:1199-1322: 7 / 12 / 8 neighbours per lattice
    particle at the 1st / 2nd / 3rd shell (self counted in the first)."""
    for case in load("lattice_shells")["cases"]:
        ng = case["ng"]
        dt = np.float32
        pos = workloads.lattice(ng, dt)
        sp = dt(1.0) / dt(ng)
        first, second, third = sp, dt(math.sqrt(2.0 * sp * sp)), dt(math.sqrt(3.0 * sp * sp))
        ghost = float(third * dt(1.1))
        wl = workloads.Workload("lattice", pos, (0.0,) * 3, (1.0,) * 3, (1, 1, 1), ghost, ghost)
        allp, _ = _with_ghosts(wl, dt)
        r = dt(third + third * dt(0.05))
        geom = ora.CellGeom(wl.box_lo, wl.box_hi, ghost, float(r), 1.013, dt)
        off, nbr = ora.verlet(allp, wl.n, geom, float(r))
        f1 = dt(first + first * dt(0.05))
        f2 = dt(second + second * dt(0.05))
        for p in range(0, wl.n, 7):
            row = nbr[off[p]:off[p + 1]]
            d = np.sqrt(((allp[:, row] - allp[:, [p]]) ** 2).sum(axis=0))
            assert (d <= f1).sum() == case["first"]
            assert ((d > f1) & (d <= f2)).sum() == case["second"]
            assert (d > f2).sum() == case["third"]


# ----------------------------------------------------------------- physics
def _two_particles(r, dt=np.float64):
    pos = np.zeros((3, 2), dt)
    pos[:, 0] = 0.5
    pos[:, 1] = 0.5
    pos[0, 1] = 0.5 + r
    return np.ascontiguousarray(pos)


def test_lj_closed_forms():
    """F(2^(1/6) sigma) = 0 and U = -eps there; F(sigma) = 24 eps / sigma."""
    sigma, eps, rc = 0.02, 1.0, 0.05
    geom = ora.CellGeom((0, 0, 0), (1, 1, 1), rc, rc, 1.013, np.float64)
    rmin = 2 ** (1 / 6) * sigma
    F, U, nn = ora.lj(_two_particles(rmin), 2, geom, rc, sigma, eps)
    assert abs(F[0, 0]) < 1e-9 * 24 * eps / sigma and abs(U[0] + eps) < 1e-12 and nn[0] == 1
    F, U, nn = ora.lj(_two_particles(sigma), 2, geom, rc, sigma, eps)
    assert math.isclose(-F[0, 0], 24 * eps / sigma, rel_tol=1e-12)  # repulsive: pushes x_0 away (-x)
    assert math.isclose(F[1, 0], 24 * eps / sigma, rel_tol=1e-12)
    F, U, nn = ora.lj(_two_particles(rc * 1.0001), 2, geom, rc, sigma, eps)
    assert nn[0] == 0 and F[0, 0] == 0


def test_lj_walk_equals_lj_on_targets():
    """The prebuilt-cell-list LJ loop (the CPU baseline's pair timing) gives the
    pinned ofo_lj results bit for bit on any target subset."""
    wl = workloads.uniform_lj(n=20000, seed=3)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, 1.013, np.float64)
    F, U, nn = ora.lj(wl.pos, wl.n, geom, wl.r_cut, 0.02, 1.0)
    cs, ci = ora.cl_fill(wl.pos, geom)
    t = np.random.default_rng(0).choice(wl.n, 500, replace=False)
    F2, U2, nn2 = ora.lj_walk(wl.pos, t, geom, cs, ci, wl.r_cut, 0.02, 1.0)
    np.testing.assert_array_equal(F2, F[t])
    np.testing.assert_array_equal(U2, U[t])
    np.testing.assert_array_equal(nn2, nn[t])


def test_lj_newton_third_law_periodic():
    wl = workloads.uniform_lj(n=4000, seed=5)
    allp, _ = _with_ghosts(wl, np.float64)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, float(np.float32(1.013)), np.float64)
    F, U, nn = ora.lj(allp, wl.n, geom, wl.r_cut, 0.02, 1.0)
    scale = np.sqrt((F ** 2).sum(axis=1)).max()
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * scale * wl.n
    # brute-force pair count
    boff, _ = ora.brute_nn(allp, wl.n, wl.r_cut)
    np.testing.assert_array_equal(nn, np.diff(boff.astype(np.int64)) - 1)


def test_cubic_spline_closed_forms():
    h = 0.01
    assert math.isclose(ora.w3(0.0, h), 1 / (math.pi * h ** 3), rel_tol=1e-15)
    assert ora.w3(2 * h, h) == 0.0 and ora.dw3(2 * h, h) == 0.0
    # continuity at q = 1
    assert math.isclose(ora.w3(h * (1 - 1e-12), h), ora.w3(h * (1 + 1e-12), h), rel_tol=1e-9)
    # integral over R^3 == 1 (radial quadrature)
    r = np.linspace(0, 2 * h, 200001)
    w = np.array([ora.w3(x, h) for x in r[::100]])
    integ = np.trapezoid(4 * math.pi * r[::100] ** 2 * w, r[::100])
    assert abs(integ - 1) < 1e-6
    # derivative matches a finite difference
    for q in (0.3, 0.99, 1.5):
        d = (ora.w3(q * h + 1e-9, h) - ora.w3(q * h - 1e-9, h)) / 2e-9
        assert math.isclose(ora.dw3(q * h, h), d, rel_tol=1e-5)


def test_sph_partition_of_unity_lattice():
    """On an unjittered lattice, rho_i / (m / dx^3) of an interior particle is the
    lattice sum sum_k W(|k| dx) dx^3, computed here by brute force over offsets;
    it is ~1 (partition of unity) for h = 1.3 dx."""
    wl = workloads.water_column(side=16, seed=1, jitter=0.0)
    dx, h = wl.params["dx"], wl.params["h"]
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, 1.013, np.float64)
    rho = ora.sph_density(wl.pos, wl.n, geom, h, wl.params["mass"])
    k = np.arange(-3, 4)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    r = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2).ravel() * dx
    lat = sum(ora.w3(x, h) for x in r if x < 2 * h) * dx ** 3
    interior = np.all((wl.pos > 3 * dx) & (wl.pos < 0.5 - 3 * dx), axis=0)
    np.testing.assert_allclose(rho[interior] / 1000.0, lat, rtol=1e-12)
    assert abs(lat - 1) < 0.01


def test_sph_force_momentum_and_repulsion():
    wl = workloads.clustered_sph(n=6000, seed=4, dtype=np.float64)
    allp, op = _with_ghosts(wl, np.float64)
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, 1.013, np.float64)
    h, m = wl.params["h"], wl.params["mass"]
    rho = ora.sph_density(allp, wl.n, geom, h, m)
    rho_all = np.concatenate([rho, rho[op]])
    prs_all = ora.tait(rho_all, wl.params["rho0"], wl.params["c0"])
    F = ora.sph_force(allp, wl.n, geom, h, m, rho_all, prs_all)
    scale = np.abs(F).max()
    assert np.abs((F * m).sum(axis=0)).max() < 1e-9 * scale * m * wl.n
    # two particles with positive pressure repel
    two = _two_particles(h)
    g2 = ora.CellGeom((0, 0, 0), (1, 1, 1), 2 * h, 2 * h, 1.013, np.float64)
    F2 = ora.sph_force(two, 2, g2, h, 1.0, np.array([1.0, 1.0]), np.array([1.0, 1.0]))
    assert F2[0, 0] < 0 < F2[1, 0]


def test_sph_two_particle_closed_forms():
    """Two particles at r = h and r = 1.5 h (h = 2^-4, so q is exact), different
    densities and pressures.  Monaghan's cubic spline (J. Comput. Phys. 110,
    1994; Rep. Prog. Phys. 68, 2005) gives, with s = 1/(pi h^3):
      W(0) = s, W(h) = s/4, W(1.5h) = s/32 (0.25 * 0.5^3),
      dW/dr(h) = -0.75 s/h, dW/dr(1.5h) = -0.1875 s/h (-0.75 * 0.5^2),
    so rho_i = m (W(0) + W(r)) and the pressure force on particle 0 is
      F_0 = -m (P_0/rho_0^2 + P_1/rho_1^2) dW/dr(r) (x_0 - x_1)/r,
    repulsive for positive pressures, equal and opposite on particle 1.  A
    dropped 1/r (h != 1), P/rho in place of P/rho^2 (rho != 1) or a swapped
    sign each fail these values."""
    h, m = 0.0625, 0.7
    s = 1.0 / (math.pi * h ** 3)
    rho = np.array([1000.0, 1100.0])
    prs = np.array([3.0e4, 5.0e4])
    g2 = ora.CellGeom((0, 0, 0), (1, 1, 1), 2 * h, 2 * h, 1.013, np.float64)
    for q, w, dw in ((1.0, 0.25, -0.75), (1.5, 0.03125, -0.1875)):
        two = _two_particles(q * h)
        assert two[0, 1] - two[0, 0] == q * h  # exact separation
        d = ora.sph_density(two, 2, g2, h, m)
        np.testing.assert_allclose(d, m * s * (1.0 + w), rtol=1e-14)
        F = ora.sph_force(two, 2, g2, h, m, rho, prs)
        expect = -m * (prs[0] / rho[0] ** 2 + prs[1] / rho[1] ** 2) * (dw * s / h) * (-1.0)
        assert expect < 0  # particle 0 (lower x) is pushed towards -x
        np.testing.assert_allclose(F[0, 0], expect, rtol=1e-13)
        np.testing.assert_allclose(F[1, 0], -expect, rtol=1e-13)
        assert F[0, 1] == F[0, 2] == F[1, 1] == F[1, 2] == 0.0
    # beyond 2h: nothing
    F = ora.sph_force(_two_particles(2.0 * h), 2, g2, h, m, rho, prs)
    assert np.all(F == 0.0)


def test_tait_closed_forms():
    """Tait EOS P = B((rho/rho0)^7 - 1), B = c0^2 rho0 / 7: P(rho0) = 0,
    P(2 rho0) = 127 c0^2 rho0 / 7, P(rho0/2) = -(127/128) c0^2 rho0 / 7."""
    rho0, c0 = 1000.0, 20.0 * math.sqrt(9.81 * 0.5)
    B = c0 * c0 * rho0 / 7.0
    P = ora.tait(np.array([rho0, 2 * rho0, 0.5 * rho0]), rho0, c0)
    assert P[0] == 0.0
    np.testing.assert_allclose(P[1], 127.0 * B, rtol=1e-15)
    np.testing.assert_allclose(P[2], -127.0 / 128.0 * B, rtol=1e-15)


class _Grid:
    def __init__(self, lo, sp, divc, dt):
        self.lo = np.ascontiguousarray(lo, dt)
        self.sp = np.ascontiguousarray(sp, dt)
        self.divc = np.ascontiguousarray(divc, np.uint32)


def face_points(dt):
    """Cell faces lo + k sp of a dyadic grid (lo = 0, 1, 2; sp = 2^-3, 2^-4, 2^-2)
    and the next representable value below each face.  x - lo is exact for all
    of them (x = 0 - lo, or lo/2 <= x <= 2 lo), and dividing by a power of two
    is exact, so the cell of every point follows from the half-open rule alone:
    a point on face k is in cell k + 1 (one padding layer below cell 1), a point
    just below it in cell k."""
    lo = np.array([0.0, 1.0, 2.0])
    sp = np.array([0.125, 0.0625, 0.25])
    nk = np.array([8, 16, 8])
    pts, cells = [], []
    for d in range(3):
        for k in range(1, nk[d] + 1):
            face = dt(lo[d] + k * sp[d])
            for x, c in ((face, k + 1), (np.nextafter(face, dt(-np.inf)), k)):
                p = [dt(lo[e] + 0.5 * sp[e]) for e in range(3)]  # cell 1 in the other dims
                p[d] = dt(x)
                want = [1, 1, 1]
                want[d] = c
                pts.append(p)
                cells.append(want)
    return lo, sp, nk, np.ascontiguousarray(np.array(pts, dt).T), np.array(cells)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_cell_key_half_open_faces(dt):
    """The cell key floor((x - lo)/sp) + 1 on exactly representable cell faces
    and one ulp below them, in the position dtype."""
    lo, sp, nk, pos, cells = face_points(dt)
    divc = nk + 3  # cells 1..nk+1 plus padding on both sides
    cid = ora.cell_ids(pos, _Grid(lo, sp, divc, dt)).astype(np.int64)
    got = np.stack([cid % divc[0], (cid // divc[0]) % divc[1], cid // (divc[0] * divc[1])], axis=1)
    np.testing.assert_array_equal(got, cells)


def test_apply_bc():
    pos = np.ascontiguousarray(np.array([[1.25, -0.25, 0.5, 1.0], [0.5, 0.5, 0.5, 0.0], [0.1, 0.2, 0.3, 0.4]]))
    out = ora.apply_bc(pos, [0, 0, 0], [1, 1, 1], [1, 1, 0])
    np.testing.assert_array_equal(out[0], [0.25, 0.75, 0.5, 0.0])
    np.testing.assert_array_equal(out[2], pos[2])
