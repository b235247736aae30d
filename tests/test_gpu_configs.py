"""The batching study (256 periodic LJ boxes) and the dynamic run (2^26
particles in power-law clusters) at their own sizes on the GPU.

Batching: the ensemble of workloads.ensemble() (256 boxes, 1,024-262,144
particles each, their own cut-offs) against the one-system path bit for bit on
a sample of boxes (smallest, largest, four random), against the oracle's FP64
LJ on the same boxes, and every box's ghost count against the oracle's.

Dynamic run: the 2^26-particle workload through the ensemble engine as the
benchmark runs it: the cell list equals the stable sort of the oracle's cell
ids, sampled Verlet rows equal a brute-force search of their stencil cells,
sampled forces match the oracle's FP64 LJ, and the shifted energy is kept over
velocity-Verlet steps that cross Verlet rebuilds."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import _lib as L
from openfpm_pdata_amd import ensemble as E
from openfpm_pdata_amd import vector_dist as vdm
from openfpm_pdata_amd import workloads
from test_gpu_ensemble import SKIN, SUB, energy, single_reference

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


class _Grid:
    """Cell geometry for the oracle's cell functions (lo, spacing, padded dims)."""

    def __init__(self, lo, sp, divc, dt):
        self.lo = np.ascontiguousarray(lo, dt)
        self.sp = np.ascontiguousarray(sp, dt)
        self.divc = np.ascontiguousarray(divc, np.uint32)


def _norm_err(got, ref):
    d = np.linalg.norm(got - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    e = d / np.maximum(mag, np.sqrt((mag ** 2).mean()))
    return e.max(), np.sqrt((e ** 2).mean())


def test_batching_study_256_boxes():
    wls = workloads.ensemble()
    systems = [E.System(wl.n, wl.r_cut, wl.params["sigma"], 1.0, 1.0, tuple(wl.box_lo), tuple(wl.box_hi))
               for wl in wls]
    pos = [wl.pos for wl in wls]
    ens = E.Ensemble(systems, torch.float32, E.Options(skin=SKIN, sub=SUB))
    ens.set(torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda())
    ens.rebuild()
    F = {prec: ens.forces(prec).cpu().numpy() for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C)}
    st = ens.status()
    print(f"batching study: {len(wls)} boxes, {ens.N} particles, {st['n_ghost']} ghosts, "
          f"{st['n_entries']} Verlet entries", flush=True)
    assert st["overflow"] == 0 and np.isfinite(F[L.OFP_PREC_F32C]).all()

    # every box's ghost count against the oracle's count/scan
    a = ens.arrays()
    N, G = ens.N, int(st["n_ghost"])
    sid = vdm.device_view(a.sid, N + G, torch.int32).cpu().numpy()
    got = np.bincount(sid[N:], minlength=len(wls))
    for s, (sy, wl) in enumerate(zip(systems, wls)):
        rv = sy.r_cut * (1.0 + SKIN)
        blo, bhi, sv = ora.local_ghost_boxes(sy.box_lo, sy.box_hi, [-rv] * 3, [rv] * 3, (1, 1, 1), np.float32)
        _, tot = ora.ghost_count(np.ascontiguousarray(wl.pos), wl.n, blo, bhi, sv)
        assert got[s] == tot, (s, got[s], tot)

    # sampled boxes: bit-identical to the one-system path, FP64 form against the oracle
    sizes = np.array([sy.n for sy in systems])
    rng = np.random.default_rng(4)
    sample = [int(np.argmin(sizes)), int(np.argmax(sizes))]
    sample += [int(b) for b in rng.choice(np.setdiff1d(np.arange(len(wls)), sample), 4, replace=False)]
    for s in sample:
        sy = systems[s]
        f0 = int(ens.first[s])
        ref = single_reference(sy, pos[s])
        for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
            np.testing.assert_array_equal(F[prec][:, f0:f0 + sy.n], ref[prec])
        vd = vdm.VectorDistGPU(sy.n, sy.box_lo, sy.box_hi, (1, 1, 1), sy.r_cut, dtype=torch.float32,
                               capacity=sy.n * 3 + 64)
        vd.pos[:, : sy.n] = torch.from_numpy(pos[s]).cuda()
        vd.ghost_get()
        allpos = np.ascontiguousarray(vd.getPosVector().cpu().numpy())
        geom = ora.CellGeom(sy.box_lo, sy.box_hi, sy.r_cut, sy.r_cut, vdm.GHOST_ENLARGE_CL, np.float32)
        oF, _, _ = ora.lj(allpos, sy.n, geom, sy.r_cut, sy.sigma, sy.eps)
        emax, _ = _norm_err(F[L.OFP_PREC_F64][:, f0:f0 + sy.n].T.astype(np.float64), oF)
        emax32, erms32 = _norm_err(F[L.OFP_PREC_F32C][:, f0:f0 + sy.n].T.astype(np.float64), oF)
        print(f"box {s}: n {sy.n} r_c {sy.r_cut:.4f}: FP64 vs oracle {emax:.2e}, FP32C vs oracle max {emax32:.2e} "
              f"rms {erms32:.2e}", flush=True)
        assert emax < 1e-6  # FP64 arithmetic, forces stored in the ensemble dtype (float)
        assert emax32 < 1e-6 and erms32 < 1e-7
        del vd


def test_dynamic_run_full_size():
    n = 1 << 26
    wl, pos, vel, rounds, n_close = workloads.power_law_clusters_lj(n=n)
    print(f"dynamic run workload: {n} particles, {rounds} redraw rounds, {n_close} close pairs", flush=True)
    assert n_close == 0
    sy = E.System(n, wl.r_cut, wl.params["sigma"], 1.0, 1.0, tuple(wl.box_lo), tuple(wl.box_hi))
    skin = 0.1
    ens = E.Ensemble([sy], torch.float32, E.Options(skin=skin, sub=1, lanes=16))
    ens.set(pos, vel)
    ens.rebuild()
    torch.cuda.synchronize()
    st = ens.status()
    N, G = ens.N, int(st["n_ghost"])
    a = ens.arrays()
    P = vdm.device_view(a.pos, 3 * int(a.cap), torch.float32).view(3, -1)[:, : N + G].cpu().numpy()
    P = np.ascontiguousarray(P)
    cv, vv = ens.lists()
    cs = vdm.device_view(cv.cell_start, int(cv.n_cells) + 1, torch.int64).cpu().numpy() & 0xFFFFFFFF
    s2n = vdm.device_view(cv.sort_to_nonsort, N + G, torch.int32).cpu().numpy().astype(np.int64)
    dom = vdm.device_view(cv.dom_sort_ids, N, torch.int32).cpu().numpy().astype(np.int64)
    _, divc, lo, sp = ens.cells()
    grid = _Grid(lo[0], sp[0], divc[0], np.float32)
    cid = ora.cell_ids(P, grid).astype(np.int64)
    print(f"cell list: {int(cv.n_cells)} cells, {N + G} particles", flush=True)
    # (1) canonical cell order: the stable sort of the oracle's cell ids
    np.testing.assert_array_equal(s2n, np.argsort(cid, kind="stable"))
    np.testing.assert_array_equal(np.diff(cs), np.bincount(cid, minlength=int(cv.n_cells)))
    assert len(dom) == N and np.all(np.diff(dom) > 0) and np.all(s2n[dom] < N)

    # (2) sampled Verlet rows == brute force over their 27 stencil cells (FP32,
    # squares summed unfused, strict r_v^2)
    off = vdm.device_view(vv.offsets, N + 1, torch.int64)
    nbr = vdm.device_view(vv.nbr, int(off[-1].item()), torch.int32)
    r2v = np.float32(sy.r_cut * (1.0 + skin)) * np.float32(sy.r_cut * (1.0 + skin))
    d0, d1 = int(divc[0][0]), int(divc[0][1])
    rng = np.random.default_rng(11)
    rows = rng.choice(N, 1500, replace=False)
    Ps = P[:, s2n]  # positions in sorted order
    cid_s = cid[s2n]
    for k in rows:
        p = dom[k]
        c = cid_s[p]
        cand = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    q = c + dx + d0 * (dy + d1 * dz)
                    cand.append(np.arange(cs[q], cs[q + 1]))
        cand = np.concatenate(cand)
        dd = Ps[:, cand] - Ps[:, [p]]
        d2 = (dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]
        want = np.sort(cand[d2 < r2v])
        o0, o1 = int(off[k]), int(off[k + 1])
        row = nbr[o0:o1].cpu().numpy().astype(np.int64)
        assert np.all(np.diff(row) > 0)
        np.testing.assert_array_equal(row, want)

    # (3) forces of sampled targets against the oracle's FP64 LJ
    gid = ens.gid().cpu().numpy().astype(np.int64)
    targets = rng.choice(N, 400, replace=False)  # storage indices
    rest = np.setdiff1d(np.arange(N + G), targets)
    allp = np.ascontiguousarray(np.concatenate([P[:, targets], P[:, rest]], axis=1))
    oF, oU, onn = ora.lj(allp, len(targets), _Grid(lo[0], sp[0], divc[0], np.float32), sy.r_cut, sy.sigma,
                         sy.eps)
    assert onn.min() >= 0 and onn.max() > 0
    for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
        F, U = ens.forces(prec, energy=True)
        got = F.cpu().numpy()[:, gid[targets]].T.astype(np.float64)
        emax, erms = _norm_err(got, oF)
        print(f"dynamic run forces, precision {prec}: max {emax:.2e} rms {erms:.2e} vs the FP64 oracle", flush=True)
        assert emax < 1e-6 and erms < 1e-7
        # per-particle energies (stored in the ensemble dtype)
        np.testing.assert_allclose(U.cpu().numpy()[gid[targets]], oU, rtol=1e-5, atol=1e-6 * np.abs(oU).max())
    del P, Ps, allp, cid, cid_s

    # (4) velocity Verlet across rebuilds: dt = 1e-3 (the fastest particles cross
    # skin r_c / 2 = 0.125 sigma within ~20 steps), shifted energy kept
    e0 = energy(ens, L.OFP_PREC_F32C)
    _, v, _ = ens.get()
    ek0 = 0.5 * float((v.double() ** 2).sum())
    r0 = ens.status()["rebuilds"]
    ens.run(L.OFP_PREC_F32C, 1e-3, 50)
    e1 = energy(ens, L.OFP_PREC_F32C)
    st = ens.status()
    print(f"dynamic run 2^26, 50 steps at dt 1e-3: {st['rebuilds'] - r0} rebuilds, shifted energy drift "
          f"{abs(e1 - e0) / ek0:.2e} of Ekin", flush=True)
    assert st["rebuilds"] - r0 >= 1 and st["overflow"] == 0
    assert abs(e1 - e0) < 1e-3 * ek0
