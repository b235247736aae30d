"""The LDS-staged cell-tile pair pass (no Verlet list), and the ensemble's
Verlet-row pass at one and at sixteen lanes per row.

One system: the cell-list pair pass on stencils of 1, 2 and 3 cells at
densities where the target items are single dense cells (split at 64 targets)
or runs of up to 8 dilute cells (the widest separations the FP32C records
meet), against the oracle's cell-list LJ: neighbour counts exact in both
forms, FP64 forces to 1e-12, FP32C to the 1e-6 / 1e-7 bound.

Ensembles (pair_walk = 1): the persistent pass over the (system, cell tile)
queue of the last rebuild's cells gives the Verlet-row pass's neighbour
counts exactly and its forces to rounding; graph replay equals direct
launches bit for bit over rebuilds; the shifted energy is kept.

Verlet rows at one lane per row (the dilute workloads' tile shape) against
sixteen lanes per row: the same pairs, forces to rounding."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import _lib as L
from openfpm_pdata_amd import ensemble as E
from openfpm_pdata_amd import vector_dist as vdm
from openfpm_pdata_amd import workloads
from test_gpu_ensemble import energy, lattice_systems, small_systems
from test_gpu_parity import make_vd, norm_err

pytestmark = pytest.mark.gpu


def cell_walk_lj(wl, dt, precision, box):
    """LJ through the cell-tile pass on cells of r_c / box with a stencil of box."""
    vd = make_vd(wl, dt)
    vd.ghost_get()
    allpos = np.ascontiguousarray(vd.getPosVector().cpu().numpy())
    opt = vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION | vdm.CL_QUANT
    cl = vd.getCellListGPU(wl.r_cut / box, opt, NNIteratorBox=box)
    vd.updateCellListGPU(cl)
    odt = torch.float64 if precision == L.OFP_PREC_F64 else torch.float32
    F = torch.zeros((3, vd.cap), dtype=odt, device="cuda")
    U = torch.zeros(vd.cap, dtype=odt, device="cuda")
    nn = torch.zeros(vd.cap, dtype=torch.int32, device="cuda")
    prm = L.LJParams(wl.r_cut, wl.params["sigma"], wl.params["eps"])
    fp = vdm.planes_of(F)
    print(f"cell tiles: n {wl.n} box {box} precision {precision}", flush=True)
    L.call("ofp_pair_lj", cl.h, None, vd.pos.data_ptr(), vd.cap, precision, prm, fp[0][0], fp[1][0], fp[2][0],
           U.data_ptr(), nn.data_ptr(), vdm._stream())
    torch.cuda.synchronize()
    n2s = cl.arrays()["n2s"].cpu().numpy().astype(np.int64)[: wl.n]
    return (F.cpu().numpy()[:, n2s].T.astype(np.float64), U.cpu().numpy()[n2s].astype(np.float64),
            nn.cpu().numpy()[n2s], allpos)


_ORACLE = {}


def oracle_lj(wl, allpos):
    """The oracle's cell-list LJ (cells of r_c) on the domain + ghost positions;
    the same for every stencil, computed once per workload."""
    if wl.n not in _ORACLE:
        geom = ora.CellGeom(wl.box_lo, wl.box_hi, wl.ghost, wl.r_cut, vdm.GHOST_ENLARGE_CL, np.float64)
        _ORACLE[wl.n] = (allpos.copy(), ora.lj(allpos, wl.n, geom, wl.r_cut, wl.params["sigma"], wl.params["eps"]))
    pos0, res = _ORACLE[wl.n]
    np.testing.assert_array_equal(pos0, allpos)
    return res


# n = 65536 in the unit cube with r_c = 0.05: ~6 particles per r_c cell (items of
# several cells), ~0.2 per r_c / 3 cell (items of up to 8 cells); n = 2^19 gives
# ~65 per r_c cell (about half the cells split into items of 64)
@pytest.mark.parametrize("n", [65536, 1 << 19])
@pytest.mark.parametrize("box", [1, 2, 3])
@pytest.mark.parametrize("precision", [L.OFP_PREC_F64, L.OFP_PREC_F32C])
def test_cell_tiles_match_oracle(n, box, precision):
    # sigma scaled with the mean spacing, so the closest pairs stay as close in
    # units of sigma as in the 65,536-particle workload
    wl = workloads.uniform_lj(n=n, seed=5, sigma=0.02 * (65536 / n) ** (1 / 3))
    F, U, nn, allpos = cell_walk_lj(wl, np.float64, precision, box)
    oF, oU, onn = oracle_lj(wl, allpos)
    np.testing.assert_array_equal(nn, onn)
    emax, erms = norm_err(F, oF)
    print(f"cell tiles n {n} box {box} precision {precision}: max {emax:.2e} rms {erms:.2e}")
    if precision == L.OFP_PREC_F64:
        assert emax < 1e-12
        np.testing.assert_allclose(U, oU, rtol=1e-11, atol=1e-12 * np.abs(oU).max())
    else:
        assert emax < 1e-6 and erms < 1e-7


def _ensemble(systems, pos, walk, vel=None, **kw):
    ens = E.Ensemble(systems, torch.float32, E.Options(skin=kw.get("skin", 0.1), sub=kw.get("sub", 2),
                                                       pair_walk=walk))
    ens.set(torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda(),
            None if vel is None else torch.from_numpy(np.ascontiguousarray(np.concatenate(vel, axis=1))).cuda())
    return ens


@pytest.mark.parametrize("sub", [1, 2, 3])
def test_ensemble_tiles_equal_verlet_rows(sub):
    systems, pos = small_systems()
    out = {}
    for walk in (0, 1):
        ens = _ensemble(systems, pos, walk, sub=sub)
        ens.rebuild()
        st = ens.status()
        assert st["overflow"] == 0 and (st["n_entries"] == 0) == (walk == 1)
        for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
            F, U = ens.forces(prec, energy=True)
            _, nn = ens.sorted_results(prec)
            # forces and energies in caller order, counts per domain row (the
            # rows of one cell list, the same for both walks)
            out[walk, prec] = (F.cpu().numpy().astype(np.float64), U.cpu().numpy().astype(np.float64),
                               nn.cpu().numpy())
        del ens
    for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
        (Fa, Ua, na), (Fb, Ub, nb) = out[0, prec], out[1, prec]
        np.testing.assert_array_equal(na, nb)
        emax, _ = norm_err(Fb.T, Fa.T)
        assert emax < (1e-6 if prec == L.OFP_PREC_F64 else 2e-6), emax  # forces stored in FP32
    # the tile pass's FP32C against its own FP64
    emax, erms = norm_err(out[1, L.OFP_PREC_F32C][0].T, out[1, L.OFP_PREC_F64][0].T)
    assert emax < 1e-6


def test_ensemble_tiles_graph_equals_direct():
    systems, pos, vel = lattice_systems()
    a = _ensemble(systems, pos, 1, vel, skin=0.2)
    b = _ensemble(systems, pos, 1, vel, skin=0.2)
    a.run(L.OFP_PREC_F32C, 5e-4, 150, use_graph=True)
    b.run(L.OFP_PREC_F32C, 5e-4, 150, use_graph=False)
    torch.cuda.synchronize()
    sa, sb = a.status(), b.status()
    assert sa["rebuilds"] == sb["rebuilds"] and sa["rebuilds"] >= 2
    for x, y in zip(a.get(), b.get()):
        torch.testing.assert_close(x, y, rtol=0, atol=0)


def test_ensemble_tiles_energy_and_trajectory():
    """600 steps with rebuilds: the shifted energy is kept as with Verlet rows,
    and the trajectory follows the Verlet-row run's to rounding."""
    systems, pos, vel = lattice_systems()
    runs = {}
    for walk in (0, 1):
        ens = _ensemble(systems, pos, walk, vel, skin=0.2)
        e0 = energy(ens, L.OFP_PREC_F64)
        _, v, _ = ens.get()
        ek = 0.5 * float((v.double() ** 2).sum())
        ens.run(L.OFP_PREC_F64, 5e-4, 600)
        e1 = energy(ens, L.OFP_PREC_F64)
        st = ens.status()
        print(f"pair_walk {walk}: E0 {e0:.6f} E1 {e1:.6f} rebuilds {st['rebuilds']}")
        assert st["rebuilds"] >= 3 and st["overflow"] == 0
        assert abs(e1 - e0) / ek < 1e-2
        runs[walk] = ens.get()[0].double()
    assert float((runs[0] - runs[1]).abs().max()) < 1e-3


def _forces_nn(systems, pos, opts):
    ens = E.Ensemble(systems, torch.float32, opts)
    ens.set(pos if torch.is_tensor(pos) else
            torch.from_numpy(np.ascontiguousarray(np.concatenate(pos, axis=1))).cuda())
    ens.rebuild()
    out = {}
    for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
        F, U = ens.forces(prec, energy=True)
        _, nn = ens.sorted_results(prec)
        out[prec] = (F.cpu().numpy().astype(np.float64), U.cpu().numpy().astype(np.float64), nn.cpu().numpy())
    return out


@pytest.mark.parametrize("sub", [1, 2, 3])
def test_row_lanes_agree(sub):
    """One Verlet row per lane (lanes = 1) against 16 lanes per row: the same
    pairs (neighbour counts exact) and the same forces to rounding (the partial
    sums are added in another order), both precisions."""
    systems, pos = small_systems()
    a = _forces_nn(systems, pos, E.Options(skin=0.1, sub=sub, lanes=1))
    b = _forces_nn(systems, pos, E.Options(skin=0.1, sub=sub, lanes=16))
    for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
        np.testing.assert_array_equal(a[prec][2], b[prec][2])
        emax, _ = norm_err(a[prec][0].T, b[prec][0].T)
        assert emax < 2e-6, (prec, emax)  # forces stored in FP32


def test_row_lanes_agree_clustered_workload():
    """The dynamic run's clustered workload at 2^18 particles (dense and dilute
    cells): one lane per row against 16 lanes per row."""
    wl, pos, vel, rounds, n_close = workloads.power_law_clusters_lj(n=1 << 18, max_rounds=80)
    sy = E.System(wl.params["n"], wl.r_cut, wl.params["sigma"], 1.0, 1.0, tuple(wl.box_lo), tuple(wl.box_hi))
    a = _forces_nn([sy], pos, E.Options(skin=0.05, sub=1, lanes=1))
    b = _forces_nn([sy], pos, E.Options(skin=0.05, sub=1, lanes=16))
    for prec in (L.OFP_PREC_F64, L.OFP_PREC_F32C):
        np.testing.assert_array_equal(a[prec][2], b[prec][2])
        emax, _ = norm_err(a[prec][0].T, b[prec][0].T)
        print(f"rows at 1 vs 16 lanes, 2^18 clustered, precision {prec}: max force difference {emax:.2e}",
              flush=True)
        assert emax < 2e-6
