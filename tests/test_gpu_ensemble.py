"""Ensembles (ofp_ens_*) against the one-system path, system by system.

Per system the ensemble must reproduce the one-system results exactly: the
same ghosts (positions, order), the same canonical cell contents, the same
Verlet rows, and bit-identical LJ forces in both precisions (same arithmetic,
same pair order).  The dynamics are checked for graph replay == direct launch
(bit-identical trajectories), rebuild bookkeeping, and energy conservation.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import _lib as L
from openfpm_pdata_amd import ensemble as E
from openfpm_pdata_amd import vector_dist as vdm

pytestmark = pytest.mark.gpu

SKIN, SUB = 0.1, 2


def small_systems(seed=7):
    rng = np.random.default_rng(seed)
    spec = [(700, 0.12, (0, 0, 0), (1, 1, 1)), (2000, 0.2, (0, 0, 0), (1, 1, 1)),
            (3500, 0.1, (-0.5, 0.25, 0.0), (0.5, 1.25, 0.75)), (1200, 0.15, (0, 0, 0), (1, 1, 1)),
            (5000, 0.09, (0, 0, 0), (1.2, 1.2, 1.2)), (1, 0.1, (0, 0, 0), (1, 1, 1))]
    systems, pos = [], []
    for n, rc, lo, hi in spec:
        lo, hi = np.array(lo, float), np.array(hi, float)
        p = (lo[:, None] + rng.random((3, n)) * (hi - lo)[:, None]).astype(np.float32)
        p = np.where(p >= hi[:, None].astype(np.float32), lo[:, None].astype(np.float32), p)
        systems.append(E.System(n, rc, 0.4 * rc, 1.0, 1.0, tuple(lo), tuple(hi)))
        pos.append(p)
    return systems, pos


def single_reference(sy, p):
    """The one-system path on system sy: ghosts, Verlet cell list and list, LJ
    forces (FP64 and FP32C) in the original order."""
    rv = sy.r_cut * (1.0 + SKIN)
    vd = vdm.VectorDistGPU(sy.n, sy.box_lo, sy.box_hi, (1, 1, 1), rv, dtype=torch.float32,
                           capacity=max(64, sy.n * 8))
    vd.pos[:, : sy.n] = torch.from_numpy(p).cuda()
    G = vd.ghost_get()
    allpos = vd.getPosVector().cpu().numpy().copy()
    opt = vdm.CL_NON_SYMMETRIC | vdm.CL_QUANT
    cl, vl = vd.getVerletGPU(rv, opt, sub=SUB)
    arr = cl.arrays()
    off, nbr = vl.arrays()
    out = dict(G=G, ghosts=allpos[:, sy.n:], cell_start=arr["cell_start"].cpu().numpy(),
               s2n=arr["s2n"].cpu().numpy(), dom=arr["dom"].cpu().numpy(), off=off.cpu().numpy(),
               nbr=nbr.cpu().numpy())
    n2s = arr["n2s"].cpu().numpy().astype(np.int64)[: sy.n]
    prm = L.LJParams(sy.r_cut, sy.sigma, sy.eps)
    for prec, tdt in ((L.OFP_PREC_F64, torch.float64), (L.OFP_PREC_F32C, torch.float32)):
        F = torch.zeros((3, vd.cap), dtype=tdt, device="cuda")
        fp = vdm.planes_of(F)
        L.call("ofp_pair_lj", cl.h, vl.h, vd.pos.data_ptr(), vd.cap, prec, prm, fp[0][0], fp[1][0], fp[2][0],
               None, None, vdm._stream())
        out[prec] = F.cpu().numpy()[:, n2s].astype(np.float32)
    return out


@pytest.fixture(scope="module")
def built():
    systems, pos = small_systems()
    ens = E.Ensemble(systems, torch.float32, E.Options(skin=SKIN, sub=SUB))
    ens.set(torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda())
    ens.rebuild()
    F = {prec: ens.forces(prec).cpu().numpy() for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C)}
    torch.cuda.synchronize()
    return systems, pos, ens, F


def test_ensemble_lists_match_single_systems(built):
    systems, pos, ens, _ = built
    st = ens.status()
    la = ens.list_arrays()
    a = ens.arrays()
    N = ens.N
    sid = vdm.device_view(a.sid, la["n_part"], torch.int32).cpu().numpy()
    cbase, _, _, _ = ens.cells()
    cs = la["cell_start"] & 0xFFFFFFFF
    G_tot = 0
    for s, sy in enumerate(systems):
        ref = single_reference(sy, pos[s])
        f0 = int(ens.first[s])
        # ghosts of system s: stored after N, ordered by source particle
        gidx = np.nonzero(sid[N:] == s)[0] + N
        assert len(gidx) == ref["G"]
        G_tot += len(gidx)
        if len(gidx):
            assert np.all(np.diff(gidx) == 1)
            gp = vdm.device_view(a.pos, 3 * a.cap, torch.float32).cpu().numpy().reshape(3, -1)[:, gidx]
            np.testing.assert_array_equal(gp, ref["ghosts"])
        g0 = int(gidx[0]) if len(gidx) else N
        # canonical cell contents, mapped to the system's own indices
        lo, hi = int(cs[cbase[s]]), int(cs[cbase[s + 1]])
        glob = la["s2n"][lo:hi].astype(np.int64)
        loc = np.where(glob < N, glob - f0, glob - g0 + sy.n)
        np.testing.assert_array_equal(loc, ref["s2n"])
        np.testing.assert_array_equal(cs[cbase[s]:cbase[s + 1] + 1] - lo, ref["cell_start"])
        # Verlet rows f0 .. f0 + n (domain-row order is system-major)
        np.testing.assert_array_equal(la["dom"][f0:f0 + sy.n].astype(np.int64) - lo, ref["dom"])
        r0, r1 = la["offsets"][f0], la["offsets"][f0 + sy.n]
        np.testing.assert_array_equal(la["offsets"][f0:f0 + sy.n + 1] - r0, ref["off"])
        np.testing.assert_array_equal(la["nbr"][r0:r1].astype(np.int64) - lo, ref["nbr"])
    assert st["n_ghost"] == G_tot and st["overflow"] == 0
    assert st["n_entries"] == la["offsets"][-1]


def test_ensemble_forces_bit_identical(built):
    systems, pos, ens, F = built
    for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
        assert np.isfinite(F[prec]).all()
    for s, sy in enumerate(systems):
        ref = single_reference(sy, pos[s])
        f0 = int(ens.first[s])
        for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
            assert np.isfinite(ref[prec]).all()
            np.testing.assert_array_equal(F[prec][:, f0:f0 + sy.n], ref[prec])
        # the FP32C form within 1e-6 (relative, normalised as in the bench) of the FP64 form
        a, b = F[L.OFP_PREC_F32C][:, f0:f0 + sy.n].astype(np.float64), F[L.OFP_PREC_F64][:, f0:f0 + sy.n]
        if sy.n > 1:
            mag = np.linalg.norm(b, axis=0)
            e = np.linalg.norm(a - b, axis=0) / np.maximum(mag, np.sqrt((mag ** 2).mean()))
            assert e.max() < 1e-6


def test_ensemble_verlet_kernels_agree():
    """The item kernels (targets in the lanes) and the thread-per-row kernels
    build the same lists, entry for entry, and the same forces."""
    systems, pos = small_systems()
    out = []
    for rows in (1, 2):
        ens = E.Ensemble(systems, torch.float32, E.Options(skin=SKIN, sub=SUB, verlet_rows=rows))
        ens.set(torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda())
        ens.rebuild()
        la = ens.list_arrays()
        out.append((la["offsets"], la["nbr"], ens.forces(L.OFP_PREC_F32C).cpu().numpy()))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_ensemble_verlet_rows_dense_agree():
    """Rows with more candidates than the thread-per-row count keeps hit masks
    for (512: cells of r_v holding ~27 particles, 27 cells a row) are finished
    by testing the rest again in the fill: the lists equal the item kernels'
    entry for entry, from a sizing build and from a second build into the
    existing capacity, beside a dilute system in the same ensemble."""
    rng = np.random.default_rng(11)
    spec = [(20000, 0.1), (900, 0.08)]
    systems, pos = [], []
    for n, rc in spec:
        systems.append(E.System(n, rc, 0.4 * rc, 1.0, 1.0, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
        p = rng.random((3, n)).astype(np.float32)
        pos.append(np.where(p >= np.float32(1.0), np.float32(0.0), p))
    out = []
    for rows in (1, 2):
        ens = E.Ensemble(systems, torch.float32, E.Options(skin=SKIN, sub=1, verlet_rows=rows))
        ens.set(torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda())
        ens.rebuild()
        la = ens.list_arrays()
        first = (la["offsets"].copy(), la["nbr"].copy())
        ens.rebuild(sizing=False)  # storage now in the first build's order: a second build
        la = ens.list_arrays()
        out.append(first + (la["offsets"].copy(), la["nbr"].copy(), ens.forces(L.OFP_PREC_F32C).cpu().numpy()))
    lens = np.diff(out[0][0].astype(np.int64))
    assert lens[:20000].mean() > 30  # dense rows: ~4/3 pi 0.11^3 20000 = 111 entries
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_ensemble_neighbour_counts_fp32c_equal_fp64(built):
    """Both precisions take the same pairs inside r_c: FP32C decides cut-off
    ties on the stored positions, as the FP64 form does."""
    _, _, ens, _ = built
    nn = {}
    for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
        ens.forces(prec, energy=True)
        nn[prec] = ens.sorted_results(prec)[1].cpu().numpy()
    assert nn[L.OFP_PREC_F64].sum() > 0
    np.testing.assert_array_equal(nn[L.OFP_PREC_F32C], nn[L.OFP_PREC_F64])


def test_ensemble_fp64_forces_match_oracle(built):
    """One system against the oracle's LJ (cell-list form, FP64)."""
    systems, pos, ens, F = built
    s = 1
    sy = systems[s]
    f0 = int(ens.first[s])
    vd = vdm.VectorDistGPU(sy.n, sy.box_lo, sy.box_hi, (1, 1, 1), sy.r_cut, dtype=torch.float32,
                           capacity=sy.n * 8)
    vd.pos[:, : sy.n] = torch.from_numpy(pos[s]).cuda()
    vd.ghost_get()
    allpos = vd.getPosVector().cpu().numpy()  # float32: the oracle bins in the position dtype
    geom = ora.CellGeom(sy.box_lo, sy.box_hi, sy.r_cut, sy.r_cut, vdm.GHOST_ENLARGE_CL, np.float32)
    oF, _, _ = ora.lj(np.ascontiguousarray(allpos), sy.n, geom, sy.r_cut, sy.sigma, sy.eps)
    got = F[L.OFP_PREC_F64][:, f0:f0 + sy.n].T.astype(np.float64)
    d = np.linalg.norm(got - oF, axis=1)
    mag = np.linalg.norm(oF, axis=1)
    e = d / np.maximum(mag, np.sqrt((mag ** 2).mean()))
    assert e.max() < 1e-6  # forces stored in the ensemble dtype (float)


def lattice_systems(seed=3):
    """Gentle LJ systems for the dynamics: jittered simple-cubic lattices with
    spacing 1.3 sigma, small velocities."""
    rng = np.random.default_rng(seed)
    systems, pos, vel = [], [], []
    for k, L_ in ((10, 1.0), (12, 1.0), (8, 0.8)):
        a = L_ / k
        sig = a / 1.3
        rc = 2.5 * sig
        g = (np.arange(k) + 0.5) * a
        z, y, x = np.meshgrid(g, g, g, indexing="ij")
        p = np.stack([x.ravel(), y.ravel(), z.ravel()]) + rng.uniform(-0.05, 0.05, (3, k ** 3)) * a
        systems.append(E.System(k ** 3, rc, sig, 1.0, 1.0, (0.0,) * 3, (L_,) * 3))
        pos.append(p.astype(np.float32))
        v = rng.normal(0, 0.3, (3, k ** 3))
        v -= v.mean(axis=1, keepdims=True)
        vel.append(v.astype(np.float32))
    return systems, pos, vel


def make_dyn(prec_opts=None):
    systems, pos, vel = lattice_systems()
    ens = E.Ensemble(systems, torch.float32, prec_opts or E.Options(skin=0.2, sub=2))
    ens.set(torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda(),
            torch.from_numpy(np.ascontiguousarray(np.concatenate(vel, axis=1))).cuda())
    return ens


U_CUT = 4.0 * (0.4 ** 12 - 0.4 ** 6)  # u(r_c) / eps for sigma = r_c / 2.5


def energy(ens, prec):
    """Kinetic + potential energy, the potential shifted by -u(r_c) per pair
    within r_c: the truncated LJ energy jumps by u(r_c) whenever a pair crosses
    r_c; the shifted sum is continuous there and is what velocity Verlet with
    the truncated force conserves."""
    F, U = ens.forces(prec, energy=True)
    _, nn = ens.sorted_results(prec)
    _, v, _ = ens.get()
    pairs = float(nn.double().sum())  # ordered pairs
    return 0.5 * float((v.double() ** 2).sum()) + 0.5 * float(U.double().sum()) - 0.5 * U_CUT * pairs


@pytest.mark.parametrize("prec", [L.OFP_PREC_F32C, L.OFP_PREC_F64])
def test_ensemble_graph_equals_direct(prec):
    a, b = make_dyn(), make_dyn()
    a.run(prec, 5e-4, 300, use_graph=True)
    b.run(prec, 5e-4, 300, use_graph=False)
    torch.cuda.synchronize()
    sa, sb = a.status(), b.status()
    # three rebuilds at least: the stored particle arrays swap at each, so the
    # third replays the first one's graph
    assert sa["rebuilds"] == sb["rebuilds"] and sa["rebuilds"] >= 3
    assert sa["recaptures"] >= 2 and sb["recaptures"] == 0
    for x, y in zip(a.get(), b.get()):
        torch.testing.assert_close(x, y, rtol=0, atol=0)


@pytest.mark.parametrize("prec", [L.OFP_PREC_F32C, L.OFP_PREC_F64])
@pytest.mark.parametrize("lanes", [1, 16])
@pytest.mark.parametrize("n_sys", [1, 3])
def test_ensemble_fused_kick_equals_split(prec, lanes, n_sys):
    """A run's steps with the kick inside the pair kernel (the default) against
    stored forces and the separate kick kernel: the same trajectory and the same
    rebuilds, bit for bit, over 150 steps (one system: its constants for every
    particle; three: each row's own)."""
    systems, pos, vel = lattice_systems()
    systems, pos, vel = systems[:n_sys], pos[:n_sys], vel[:n_sys]
    out = []
    for split in (0, 1):
        ens = E.Ensemble(systems, torch.float32, E.Options(skin=0.2, sub=2, lanes=lanes, split_kick=split))
        ens.set(torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda(),
                torch.from_numpy(np.ascontiguousarray(np.concatenate(vel, axis=1))).cuda())
        ens.run(prec, 5e-4, 150)
        torch.cuda.synchronize()
        out.append((ens.status(), [t.clone() for t in ens.get()]))
    (sa, ga), (sb, gb) = out
    assert sa["rebuilds"] == sb["rebuilds"] and sa["rebuilds"] >= 2
    for x, y in zip(ga, gb):
        torch.testing.assert_close(x, y, rtol=0, atol=0)


@pytest.mark.parametrize("prec", [L.OFP_PREC_F32C])
def test_ensemble_graph_mask_only_overflow(prec):
    """A captured rebuild whose Verlet masks no longer fit while the entries
    still do: the mask buffer is reallocated by the sizing rebuild that follows,
    so every graph captured before it is dropped and captured again (a stale
    graph would keep the freed buffer and its old capacity), and the trajectory
    equals direct launches bit for bit."""
    a, b = make_dyn(), make_dyn()
    for ens, g in ((a, True), (b, False)):
        ens.run(prec, 5e-4, 3, use_graph=g)
        # a graph-replayed rebuild with the sized mask buffer (captured once) ...
        ens.test_masks(0, pending=True)
        ens.run(prec, 5e-4, 2, use_graph=g)
    st_a0 = a.status()
    assert st_a0["recaptures"] >= 2
    for ens, g in ((a, True), (b, False)):
        # ... then a mask buffer far too small for the same configuration: the
        # next rebuild overflows the masks while its entries still fit
        old = ens.test_masks(64, pending=True)
        assert old > 64
        ens.run(prec, 5e-4, 3, use_graph=g)
        assert ens.status()["overflow"] == 0
    assert a.status()["recaptures"] > st_a0["recaptures"] and b.status()["recaptures"] == 0
    assert a.status()["rebuilds"] == b.status()["rebuilds"]
    for x, y in zip(a.get(), b.get()):
        torch.testing.assert_close(x, y, rtol=0, atol=0)


def test_ensemble_energy_conservation():
    """Velocity Verlet, dt = 5e-4 (~1/130 of the LJ time sigma sqrt(m/eps)),
    600 steps with rebuilds, FP32C forces: the shifted energy drifts by less than
    1e-2 of the initial kinetic energy (the truncated force still jumps at r_c)."""
    ens = make_dyn()
    e0 = energy(ens, L.OFP_PREC_F32C)
    _, v, _ = ens.get()
    ek = 0.5 * float((v.double() ** 2).sum())
    ens.run(L.OFP_PREC_F32C, 5e-4, 600)
    e1 = energy(ens, L.OFP_PREC_F32C)
    st = ens.status()
    print(f"ensemble energy: E0 {e0:.6f} E1 {e1:.6f} Ekin0 {ek:.3f} rebuilds {st['rebuilds']}")
    assert st["rebuilds"] >= 3 and st["overflow"] == 0
    assert abs(e1 - e0) / ek < 1e-2


def test_ensemble_rebuild_after_steps_matches_fresh_build():
    """The lists of a captured rebuild after stepping equal a sizing rebuild of an
    ensemble set to the same (stored-order) positions."""
    ens = make_dyn()
    ens.run(L.OFP_PREC_F32C, 5e-4, 40)
    ens.rebuild(sizing=False)
    la = ens.list_arrays()
    pos, _, _ = ens.get()
    fresh = make_dyn()
    fresh.set(pos)
    fresh.rebuild()
    lb = fresh.list_arrays()
    # stored orders differ (the stepped ensemble is in cell order); compare the
    # neighbour sets of every caller index
    def nsets(ens_, la_):
        gid = ens_.gid().cpu().numpy().astype(np.int64)
        N = ens_.N
        s2n = la_["s2n"].astype(np.int64)
        out = {}
        for k in range(N):
            f = la_["dom"][k]
            owner = gid[s2n[f]]
            row = la_["nbr"][la_["offsets"][k]:la_["offsets"][k + 1]].astype(np.int64)
            o = s2n[row]
            out[owner] = (len(row), tuple(sorted(gid[o[o < N]])))
        return out
    na, nb = nsets(ens, la), nsets(fresh, lb)
    assert len(na) == ens.N and na == nb


def test_dynamic_run_workload_small():
    """The dynamic run's workload at 2^16 particles (LJ units, power-law
    clusters, overlaps closer than 0.85 sigma redrawn): no close pair is left,
    the first forces are finite and bounded by the d_min repulsion, and 50
    velocity-Verlet steps at dt = 1e-3 keep the shifted energy to 1e-3 of the
    kinetic energy."""
    from openfpm_pdata_amd import workloads

    wl, pos, vel, rounds, n_close = workloads.power_law_clusters_lj(n=1 << 16, max_rounds=80)
    assert n_close == 0 and rounds < 80
    sy = E.System(wl.params["n"], wl.r_cut, wl.params["sigma"], 1.0, 1.0, tuple(wl.box_lo), tuple(wl.box_hi))
    ens = E.Ensemble([sy], torch.float32, E.Options(skin=0.1, sub=1))
    ens.set(pos, vel)
    e0 = energy(ens, L.OFP_PREC_F32C)
    F0, _ = ens.sorted_results(L.OFP_PREC_F32C)
    fmax = float(torch.linalg.vector_norm(F0, dim=0).max())
    # |F| of one pair at d_min = 0.85: 24 (2 / 0.85^13 - 1 / 0.85^7) / 0.85; a
    # particle has few neighbours that close, so a few times that bounds |F|
    f_pair = 24.0 * (2.0 / 0.85 ** 13 - 1.0 / 0.85 ** 7) / 0.85
    assert torch.isfinite(F0).all() and fmax < 8 * f_pair, fmax
    ek0 = 0.5 * float((vel.double() ** 2).sum())
    ens.run(L.OFP_PREC_F32C, 1e-3, 50)
    e1 = energy(ens, L.OFP_PREC_F32C)
    print(f"dynamic workload 2^16: redraw rounds {rounds}, |F|max {fmax:.1f}, "
          f"shifted energy drift {abs(e1 - e0) / ek0:.2e} of Ekin")
    assert abs(e1 - e0) < 1e-3 * ek0
