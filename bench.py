#!/usr/bin/env python3
"""Benchmark of the vector_dist neighbour-search hot path on one MI355X.

Workload (clustered SPH): 4,194,304 FP32 particles in the periodic unit cube,
half uniform and half from 64 Gaussian clusters (sigma 0.02, wrapped); cubic
spline SPH with h = 1.3 dx, dx = N^(-1/3), r_c = 2h; the cell list is reordered
and the Verlet list built with a 10 % skin.  One step =
    ghost_get (label + emit)
    updateCellListGPU (cell list at r_c, reorder)
    Verlet build (r_c * 1.1, own cell list, reorder)
    SPH density (Verlet) -> ghost sync of rho -> Tait EOS -> SPH force (Verlet)
    restoreOrder of the forces
run once in the FP32-compensated form (FP32 storage: quantised cell-relative
records, FP32 P/rho^2 and forces; pairs evaluated in FP64) and once in plain FP64 on an FP64
copy of the particles.  `value` = pair interactions per second of the
FP32-compensated step (ordered pairs i != j with d < r_c, counted in both the
density and the force pass) over the whole step.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=200000)
    ap.add_argument("--sub", type=int, default=2, help="Verlet cell list: cells of r/sub, stencil sub")
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with a 1 GiB device-to-device copy (calibrates HBM byte counters)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batching study")
    ap.add_argument("--no-remote", action="store_true", help="skip the remote-ghost (send side) study")
    ap.add_argument("--no-map", action="store_true", help="skip the map (migration) study")
    ap.add_argument("--no-sym", action="store_true", help="skip the symmetric-list study")
    ap.add_argument("--no-sparse", action="store_true", help="skip the sparse cell-list study")
    ap.add_argument("--cl-scale-n", type=int, default=1 << 26, help="particles of the cell-list-at-scale build (0: skip)")
    ap.add_argument("--dyn-n", type=int, default=1 << 26, help="particles of the dynamic run (0: skip)")
    ap.add_argument("--dyn-steps", type=int, default=100)
    ap.add_argument("--ens-boxes", type=int, default=256)
    ap.add_argument("--tune-budget", type=float, default=40.0, help="autotuner time budget (s)")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


class Step:
    """One pass of the hot path on one vector (FP32C or FP64)."""

    def __init__(self, wl, precision, sub=2):
        import torch

        from openfpm_pdata_amd import _lib as L
        from openfpm_pdata_amd import vector_dist as vdm

        self.L, self.vdm, self.torch = L, vdm, torch
        self.wl = wl
        self.prec = precision
        f32 = precision == L.OFP_PREC_F32C
        self.tdt = torch.float32 if f32 else torch.float64
        self.odt = torch.float32 if f32 else torch.float64
        p = wl.params
        self.r_verlet = p["r_verlet"]
        props = {"rho": (torch.float64, 1), "prs": (torch.float64, 1), "pterm": (self.odt, 1),
                 "force": (self.odt, 3)}
        self.vd = vdm.VectorDistGPU(wl.n, wl.box_lo, wl.box_hi, wl.bc, self.r_verlet, dtype=self.tdt, props=props,
                                    capacity=int(wl.n * 1.25))
        self.x0 = torch.from_numpy(wl.pos.astype(np.float32 if f32 else np.float64)).cuda()
        self.vd.pos[:, : wl.n] = self.x0
        self.prm = L.SPHParams(p["h"], p["mass"], p["rho0"], p["c0"])
        opt = vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION | vdm.CL_QUANT
        self.cl = self.vd.getCellListGPU(wl.r_cut, opt)
        self.clv = self.vd.getCellListGPU(self.r_verlet / sub, opt, NNIteratorBox=sub,
                                          ghostEnlargeFactor=vdm.GHOST_ENLARGE_VL)
        self.vl = vdm.VerletListGPU()
        self.ev = {}
        self.times = {}

    def _mark(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.ev.setdefault(name, []).append(e)

    def run(self, record=False):
        vd, L, vdm, cl, clv = self.vd, self.L, self.vdm, self.cl, self.clv
        st = vdm._stream()
        m = self._mark if record else (lambda name: None)
        m("t0")
        vd.ghost_get()
        m("ghost")
        vd.updateCellListGPU(cl)
        m("cell_list")
        vd.restoreOrder(cl)
        vd.updateCellListGPU(clv)
        self.vl.build(clv, vd.pos, self.r_verlet)
        m("verlet")
        rho, prs, pterm, F = (vd.prp[k] for k in ("rho", "prs", "pterm", "force"))
        L.call("ofp_sph_density", clv.h, self.vl.h, vd.pos.data_ptr(), vd.cap, self.prec, self.prm, rho.data_ptr(), st)
        m("density")
        vd.ghost_sync_sorted(clv, ("rho",))
        L.call("ofp_sph_eos", self.prec, rho.data_ptr(), vd.size, self.prm, prs.data_ptr(), pterm.data_ptr(), st)
        m("eos")
        fp = vdm.planes_of(F)
        L.call("ofp_sph_force", clv.h, self.vl.h, vd.pos.data_ptr(), vd.cap, self.prec, self.prm, pterm.data_ptr(),
               fp[0][0], fp[1][0], fp[2][0], st)
        m("force")
        vd.restoreOrder(clv, ("force", "rho"))
        m("end")

    def phase_ms(self):
        names = ["t0", "ghost", "cell_list", "verlet", "density", "eos", "force", "end"]
        out = {}
        for a, b in zip(names[:-1], names[1:]):
            ts = [ea.elapsed_time(eb) for ea, eb in zip(self.ev[a], self.ev[b])]
            out[b] = float(np.mean(ts))
        return out

    def n_pairs(self):
        """Ordered pairs i != j with d < r_c (domain i), from a Verlet list at r_c."""
        vdm = self.vdm
        vd = self.vd
        vd.ghost_get()
        cl, vl = vd.getVerletGPU(self.wl.r_cut)
        v = vl.view()
        pairs = int(v.n_entries) - self.wl.n  # self entries removed
        vd.restoreOrder(cl)
        del vl, cl
        return pairs, int(vd.size - vd.ghostMarker), int(self.vl.view().n_entries)


def error_vs_fp64(step):
    """FP32C results against FP64 arithmetic on the same inputs: one pass of
    the FP32C step's structures (same positions, ghosts and Verlet list) run
    once with OFP_PREC_F32C and once with OFP_PREC_F64 (positions promoted)."""
    import torch

    L, vdm = step.L, step.vdm
    vd, clv, wl = step.vd, step.clv, step.wl
    st = vdm._stream()
    vd.ghost_get()
    vd.updateCellListGPU(clv)
    step.vl.build(clv, vd.pos, step.r_verlet)
    out = {}
    for prec, pdt in ((L.OFP_PREC_F32C, torch.float32), (L.OFP_PREC_F64, torch.float64)):
        rho = torch.zeros(vd.cap, dtype=torch.float64, device="cuda")
        pterm = torch.zeros(vd.cap, dtype=pdt, device="cuda")
        F = torch.zeros((3, vd.cap), dtype=pdt, device="cuda")
        fp = vdm.planes_of(F)
        L.call("ofp_sph_density", clv.h, step.vl.h, vd.pos.data_ptr(), vd.cap, prec, step.prm, rho.data_ptr(), st)
        L.call("ofp_ghost_sync_sorted", vd._gh, clv.h, vdm.plane_array([(rho.data_ptr(), 8)]), 1, st)
        L.call("ofp_sph_eos", prec, rho.data_ptr(), vd.size, step.prm, None, pterm.data_ptr(), st)
        L.call("ofp_sph_force", clv.h, step.vl.h, vd.pos.data_ptr(), vd.cap, prec, step.prm, pterm.data_ptr(),
               fp[0][0], fp[1][0], fp[2][0], st)
        dom = clv.arrays()["dom"].long()
        out[prec] = (rho[dom].cpu().numpy(), F[:, dom].double().cpu().numpy())
    vd.restoreOrder(clv)
    (r32, F32), (r64, F64) = out[L.OFP_PREC_F32C], out[L.OFP_PREC_F64]
    d = np.sqrt(((F32 - F64) ** 2).sum(axis=0))
    mag = np.sqrt((F64 ** 2).sum(axis=0))
    e = d / np.maximum(mag, np.sqrt((mag ** 2).mean()))
    er = np.abs(r32 - r64) / np.abs(r64)
    return {"force_max": float(e.max()), "force_rms": float(np.sqrt((e ** 2).mean())),
            "rho_max": float(er.max()), "rho_rms": float(np.sqrt((er ** 2).mean())),
            "reference": "FP64 arithmetic on the same positions, ghosts and Verlet list"}


def timed_steps(step, warmup, steps):
    import torch

    for _ in range(warmup):
        step.run()
    torch.cuda.synchronize()
    step.ev.clear()
    t0 = time.perf_counter()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        step.run(record=True)
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return s.elapsed_time(e) / steps, wall


def cpu_baseline(wl, sample, rho_all_sorted=None):
    """The oracle (single thread) on a bounded sample: SPH density + force for
    `sample` random target particles of the same workload, against all of its
    particles and ghosts."""
    from oracle import oracle as ora

    g = wl.ghost
    dt = np.float32
    blo, bhi, sv = ora.local_ghost_boxes(wl.box_lo, wl.box_hi, [-g] * 3, [g] * 3, wl.bc, dt)
    sh = ora.shift_vectors(wl.box_lo, wl.box_hi, dt)
    rng = np.random.default_rng(1)
    perm = rng.permutation(wl.n)
    pos = np.ascontiguousarray(wl.pos[:, perm].astype(dt))
    gp, op, _ = ora.ghost_emit(pos, wl.n, blo, bhi, sv, sh)
    allp = np.ascontiguousarray(np.concatenate([pos, gp], axis=1))
    geom = ora.CellGeom(wl.box_lo, wl.box_hi, g, wl.r_cut, float(np.float32(1.013)), dt)
    h, mass = wl.params["h"], wl.params["mass"]
    t0 = time.perf_counter()
    rho = ora.sph_density(allp, sample, geom, h, mass)
    t1 = time.perf_counter()
    rho_all = np.full(allp.shape[1], wl.params["rho0"])
    rho_all[:sample] = rho
    prs = ora.tait(rho_all, wl.params["rho0"], wl.params["c0"])
    t2 = time.perf_counter()
    ora.sph_force(allp, sample, geom, h, mass, rho_all, prs)
    t3 = time.perf_counter()
    off, _ = ora.verlet(allp, sample, geom, wl.r_cut)
    pairs = int(off[-1]) - sample
    secs = (t1 - t0) + (t3 - t2)
    return {"value": 2 * pairs / secs, "unit": "pair interactions/s", "cores": 1, "kind": "port",
            "sample": f"oracle SPH density + force for {sample} random targets of the {wl.n}-particle "
                      f"clustered workload ({pairs} pairs, {secs:.1f} s)"}


def load_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/pmc_traffic.json: FETCH_SIZE and WRITE_SIZE in KiB per dispatch
    from two rocprofv3 --pmc passes of this benchmark, each scaled by the factor
    the same counter reads for a 1 GiB copy)."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(p):
        return None
    d = json.load(open(p))
    k = d.get("kernels", {}).get(kernel)
    if not k:
        return None
    b = 1024.0 * (k["fetch_kib"] * d["fetch_factor"] + k["write_kib"] * d["write_factor"])
    return {"bytes": b, "source": "profiles/pmc_traffic.json"}


def _evt_ms(fn, reps=3):
    import torch

    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def celllist_at_scale(args):
    """updateCellListGPU on 2^26 uniform FP32 particles in the periodic unit cube
    (~9 per cell, 195^3 cells): a working set well beyond the 256 MB Infinity
    Cache.  Built once from random order, then again from the cell order the
    first build produced (the order a run that keeps its particles sorted feeds
    in).  Algorithmic bytes per build: the positions read and written in cell
    order (3s each), sort_to_nonsort and nonsort_to_sort (4 B each) per particle,
    the domain sort ids (4 B per domain particle) and the cell starts (4 B per
    cell + 1): N·(6s + 8) + 4·N_dom + 4(C+1)."""
    import torch

    from openfpm_pdata_amd import vector_dist as vdm

    n = args.cl_scale_n
    g = torch.Generator(device="cuda").manual_seed(0x0F9D0005)
    pos = torch.rand((3, n), generator=g, device="cuda", dtype=torch.float32)
    r = (9.0 / n) ** (1.0 / 3.0)
    blo, bhi, div = vdm.cl_param(torch.float32, (0.0,) * 3, (1.0,) * 3, r, [0.0] * 3, [0.0] * 3, 1.0)
    cl = vdm.CellListGPU(torch.float32, blo, bhi, div, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION, r)
    out_pos = torch.empty_like(pos)
    C = int(np.prod(np.asarray(div, np.int64) + 2))
    alg = n * (12 + 12 + 8) + 4 * n + 4 * (C + 1)
    res = {"n_particles": n, "n_cells": C, "algorithmic_bytes": alg}
    for name in ("random_order", "cell_order"):
        t = _evt_ms(lambda: cl.construct(pos, n, n, out_pos), 3)
        res[name] = {"ms": t, "mpart_s": n / (t * 1e-3) / 1e6, "achieved_gbs": alg / (t * 1e-3) / 1e9,
                     "frac_hbm": alg / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
        pos = out_pos.clone()  # next build reads the sorted positions
    del cl, pos, out_pos
    torch.cuda.empty_cache()
    return res


def remote_section(args):
    """The send side of ghost_get for one rank of the clustered workload (4,194,304
    FP32 particles) cut 2 x 2 x 2 over 8 ranks, periodic, ghost width r_c (1 + skin):
    rank 0's particles are labelled against the 26 internal ghost boxes of its 7
    neighbours, ordered by destination and packed with two property planes
    (FP64 and FP32), as ghost_get<prp...> does before its exchange.
    Algorithmic bytes: label 3s read per particle and per output 4 (processor)
    + 8 (particle | shift) written; pack per output 3s + 8 + 4 gathered and written
    and 8 read for its id."""
    import torch

    from openfpm_pdata_amd import remote as R
    from openfpm_pdata_amd import workloads

    wl = workloads.clustered_sph(n=args.n)
    ghost = wl.params["r_verlet"]
    dims = (2, 2, 2)
    dlo, dhi, blo, bhi, proc, sid, sh = R.regular_decomposition(torch.float32, wl.box_lo, wl.box_hi, wl.bc, dims, 0,
                                                                ghost)
    pos = wl.pos
    inside = np.all((pos >= dlo[:, None]) & (pos < dhi[:, None]), axis=0)
    mine = np.ascontiguousarray(pos[:, inside])
    n = mine.shape[1]
    pt = torch.from_numpy(mine).cuda().contiguous()
    g = R.GhostRemote(torch.float32, blo, bhi, proc, sid, sh)
    rho = torch.rand(n, dtype=torch.float64, device="cuda")
    prs = torch.rand(n, dtype=torch.float32, device="cuda")
    n_out, n_procs = g.label(pt, n)
    send = torch.empty((3, max(1, n_out)), dtype=torch.float32, device="cuda")
    srho = torch.empty(max(1, n_out), dtype=torch.float64, device="cuda")
    sprs = torch.empty(max(1, n_out), dtype=torch.float32, device="cuda")
    pin = [(rho.data_ptr(), 8), (prs.data_ptr(), 4)]
    pout = [(srho.data_ptr(), 8), (sprs.data_ptr(), 4)]
    t_label = _evt_ms(lambda: g.label(pt, n), 5)
    t_pack = _evt_ms(lambda: g.pack(pt, send, pin, pout), 5)
    procs, counts = g.procs()
    b_label = n * 12 + n_out * (4 + 8)
    b_pack = n_out * 2 * (12 + 8 + 4) + n_out * 8
    res = {"workload": "clustered 4194304, 2x2x2 periodic, rank 0, ghost r_c (1 + skin)", "n_rank": n,
           "n_out": n_out, "n_procs": n_procs, "per_proc": {int(p): int(c) for p, c in zip(procs, counts)},
           "label_ms": t_label, "pack_ms": t_pack,
           "label": {"algorithmic_bytes": b_label, "achieved_gbs": b_label / (t_label * 1e-3) / 1e9},
           "pack": {"algorithmic_bytes": b_pack, "achieved_gbs": b_pack / (t_pack * 1e-3) / 1e9,
                    "frac_hbm": b_pack / (t_pack * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "note": "label includes its two host read-backs (output total, per-processor counts)"}
    g.close()
    del pt, send, srho, sprs, rho, prs
    torch.cuda.empty_cache()
    return res


def map_section(args):
    """map for one rank of the clustered workload (4,194,304 FP32 particles) cut
    2 x 2 x 2 over 8 ranks, periodic: rank 0's particles, each moved by a
    uniform step of up to half the skin per component (as between two Verlet
    rebuilds), are wrapped, labelled with their processor, ordered by
    destination and split into the staying particles and the send buffer with
    two property planes (FP64 and FP32).
    Algorithmic bytes: label 3s read and an 8-byte ordered id written per
    particle; pack per particle 3s + 8 + 4 gathered and written and 8 read for
    its id."""
    import torch

    from openfpm_pdata_amd import remote as R
    from openfpm_pdata_amd import workloads

    wl = workloads.clustered_sph(n=args.n)
    dims = (2, 2, 2)
    slo, shi, sp = R.regular_subdomains(torch.float32, wl.box_lo, wl.box_hi, dims)
    pos = wl.pos
    inside = np.all((pos >= slo[0][:, None]) & (pos < shi[0][:, None]), axis=0)
    mine = pos[:, inside].astype(np.float64)
    half_skin = 0.5 * (wl.params["r_verlet"] - wl.r_cut)
    rng = np.random.default_rng(5)
    mine = np.ascontiguousarray((mine + rng.uniform(-half_skin, half_skin, mine.shape)).astype(np.float32))
    n = mine.shape[1]
    pt = torch.from_numpy(mine).cuda().contiguous()
    m = R.Map(torch.float32, slo, shi, sp, np.asarray(wl.box_lo, np.float32), np.asarray(wl.box_hi, np.float32),
              wl.bc, 0)
    rho = torch.rand(n, dtype=torch.float64, device="cuda")
    prs = torch.rand(n, dtype=torch.float32, device="cuda")
    n_stay, n_dest, n_lost = m.label(pt, n)
    stay = torch.empty((3, max(1, n_stay)), dtype=torch.float32, device="cuda")
    send = torch.empty((3, max(1, m.n_send)), dtype=torch.float32, device="cuda")
    bufs = [torch.empty(max(1, c), dtype=d, device="cuda") for c in (n_stay, m.n_send)
            for d in (torch.float64, torch.float32)]
    pin = [(rho.data_ptr(), 8), (prs.data_ptr(), 4)]
    pst = [(bufs[0].data_ptr(), 8), (bufs[1].data_ptr(), 4)]
    psd = [(bufs[2].data_ptr(), 8), (bufs[3].data_ptr(), 4)]
    t_label = _evt_ms(lambda: m.label(pt, n), 5)
    t_pack = _evt_ms(lambda: m.pack(pt, stay, send, pin, pst, psd), 5)
    procs, counts = m.procs()
    n_kept = n - n_lost
    b_label = n * (12 + 8)
    b_pack = n_kept * 2 * (12 + 8 + 4) + n_kept * 8
    res = {"workload": "clustered 4194304, 2x2x2 periodic, rank 0, moved up to skin/2", "n_rank": n,
           "n_stay": n_stay, "n_send": m.n_send, "n_lost": n_lost,
           "per_proc": {int(p): int(c) for p, c in zip(procs, counts)},
           "label_ms": t_label, "pack_ms": t_pack,
           "label": {"algorithmic_bytes": b_label, "achieved_gbs": b_label / (t_label * 1e-3) / 1e9},
           "pack": {"algorithmic_bytes": b_pack, "achieved_gbs": b_pack / (t_pack * 1e-3) / 1e9,
                    "frac_hbm": b_pack / (t_pack * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "note": "label = wrap + processor + destination order, with its host read-back of the counts"}
    m.close()
    del pt, stay, send, bufs, rho, prs
    torch.cuda.empty_cache()
    return res


def sym_section(args):
    """Symmetric Verlet list and Newton's-third-law LJ on 4,194,304 uniform FP32
    particles in the periodic unit cube, r_c = 0.015 (~59 neighbours), Verlet
    rows at 1.1 r_c (cells of r/2): the full FP64 LJ pass over the full list
    against the symmetric half (built from the full list), its LJ pass with
    atomic adds to both ends, and ghost_put<add_> of forces and counts."""
    import torch

    from openfpm_pdata_amd import _lib as L
    from openfpm_pdata_amd import vector_dist as vdm
    from openfpm_pdata_amd import workloads

    wl = workloads.uniform_lj(n=args.n, dtype=np.float32, r_cut=0.015, sigma=0.006)
    # positions on multiples of 2^-20: periodic images are exact, so a pair's
    # distance is the same from both ends and the counts must agree exactly
    q = np.round(wl.pos.astype(np.float64) * 2.0 ** 20) / 2.0 ** 20
    wl.pos[...] = np.where(q >= 1.0, q - 1.0, q).astype(np.float32)
    r = wl.r_cut * 1.1
    vd = vdm.VectorDistGPU(wl.n, wl.box_lo, wl.box_hi, wl.bc, r, dtype=torch.float32)
    vd.pos[:, : wl.n] = torch.from_numpy(np.ascontiguousarray(wl.pos)).cuda()
    vd.ghost_get()
    cl, full, sym = vd.getVerletSymGPU(r, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION, sub=2)
    sig, eps = wl.params["sigma"], wl.params["eps"]
    prm = L.LJParams(wl.r_cut, sig, eps)
    Ff = torch.zeros((3, vd.cap), dtype=torch.float64, device="cuda")
    nf = torch.zeros(vd.cap, dtype=torch.int32, device="cuda")

    def full_pass():
        L.call("ofp_pair_lj", cl.h, full.h, vd.pos.data_ptr(), vd.cap, L.OFP_PREC_F64, prm, Ff[0].data_ptr(),
               Ff[1].data_ptr(), Ff[2].data_ptr(), None, nf.data_ptr(), vdm._stream())

    out = {}

    def sym_pass():
        F, _ = vdm.pair_lj_sym(cl, sym, vd.pos, wl.r_cut, sig, eps, counts=False)
        vd.ghost_put_sorted(cl, [F[0], F[1], F[2]])
        out["F"] = F

    def sym_pass_counts():
        F, nn = vdm.pair_lj_sym(cl, sym, vd.pos, wl.r_cut, sig, eps)
        vd.ghost_put_sorted(cl, [F[0], F[1], F[2], nn])
        out["F"], out["nn"] = F, nn

    t_full = _evt_ms(full_pass, 5)
    t_build = _evt_ms(lambda: sym.build_sym(full, vd._gh), 3)
    t_sym = _evt_ms(sym_pass, 5)
    t_sym_n = _evt_ms(sym_pass_counts, 3)
    torch.cuda.synchronize()
    dom = cl.arrays()["dom"].long()
    pairs = int(nf[dom].long().sum().item())  # ordered pairs within r_c (the pass skips self)
    mism = int((out["nn"][dom] != nf[dom]).sum().item())
    fr = Ff[:, dom]
    ferr = float(((out["F"][:, dom] - fr).abs().max() / fr.norm(dim=0).max()).item())
    e_full, e_sym = int(full.view().n_entries), int(sym.view().n_entries)
    res = {"workload": "uniform LJ 4194304 FP32, periodic, r_c 0.015, Verlet 1.1 r_c", "entries_full": e_full,
           "entries_sym": e_sym, "ordered_pairs": pairs, "full_lj_ms": t_full, "sym_build_ms": t_build,
           "sym_lj_put_ms": t_sym, "sym_lj_put_with_counts_ms": t_sym_n, "full_pairs_per_s": pairs / (t_full * 1e-3),
           "sym_pairs_per_s": pairs / (t_sym * 1e-3), "count_mismatches_after_put": mism,
           "force_diff_vs_full": ferr,
           "note": "pairs/s counts ordered pairs (each interaction twice) for both passes"}
    del vd, cl, full, sym, Ff, nf, out
    torch.cuda.empty_cache()
    return res


def sparse_section(args):
    """Sparse form of the step's cell list (clustered 4,194,304 FP32 particles
    plus ghosts, cells of r_c): build time of the compaction and neighbour
    table, non-empty cells against all cells."""
    import torch

    from openfpm_pdata_amd import _lib as L
    from openfpm_pdata_amd import vector_dist as vdm
    from openfpm_pdata_amd import workloads

    wl = workloads.clustered_sph(n=args.n)
    vd = vdm.VectorDistGPU(wl.n, wl.box_lo, wl.box_hi, wl.bc, wl.ghost, dtype=torch.float32)
    vd.pos[:, : wl.n] = torch.from_numpy(np.ascontiguousarray(wl.pos)).cuda()
    vd.ghost_get()
    cl = vd.getCellListGPU(wl.r_cut, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION)
    vd.updateCellListGPU(cl)
    t = _evt_ms(lambda: L.call("ofp_cl_build_sparse", cl.h, vdm._stream()), 5)
    sp = cl.sparse()
    v = cl.view()
    n_nz = int(sp["key"].numel())
    res = {"workload": "clustered 4194304 FP32 + ghosts, cells of r_c", "cells": int(v.n_cells),
           "nonempty_cells": n_nz, "build_ms": t,
           "bytes_dense_starts": 8 * (int(v.n_cells) + 1), "bytes_sparse": 4 * (2 * n_nz + 1) + 4 * 27 * n_nz,
           "note": "build_ms includes the read-back of the non-empty count"}
    del vd, cl, sp
    torch.cuda.empty_cache()
    return res


def dynamic_section(args):
    """Dynamic run in Lennard-Jones units (sigma = eps = m = 1): 2^26 FP32
    particles from 512 Gaussian clusters of weights k^-1.5 in a periodic box of
    mean density 0.05 (cluster centres 0.4), r_c = 2.5, Maxwell velocities at
    T = 1, velocity Verlet at dt = 1e-4 for 100 steps, Verlet lists (skin 10 %)
    rebuilt when a particle may have moved beyond skin r_c / 2, ghosts refreshed
    every step; each step one hipGraph replay.  Initial overlaps closer than
    0.85 sigma are removed by redrawing (untimed)."""
    import torch

    from openfpm_pdata_amd import _lib as L
    from openfpm_pdata_amd import ensemble as E
    from openfpm_pdata_amd import workloads

    t_init = time.perf_counter()
    wl, pos, vel, rounds, n_close = workloads.power_law_clusters_lj(n=args.dyn_n)
    t_init = time.perf_counter() - t_init
    sy = E.System(wl.params["n"], wl.r_cut, wl.params["sigma"], 1.0, 1.0, tuple(wl.box_lo), tuple(wl.box_hi))
    ens = E.Ensemble([sy], torch.float32, E.Options(skin=0.1, sub=1, lanes=16, waves_per_cu=32))
    ens.set(pos, vel)
    del pos, vel
    torch.cuda.empty_cache()
    u_cut = 4.0 * (0.4 ** 12 - 0.4 ** 6)  # u(r_c) at r_c = 2.5 sigma

    def shifted_energy():
        """Kinetic + LJ energy with u(r_c) taken off every pair inside r_c (the
        quantity velocity Verlet with the truncated force conserves), FP64 sums."""
        _, U = ens.forces(L.OFP_PREC_F32C, energy=True)
        _, nn = ens.sorted_results(L.OFP_PREC_F32C)
        _, v, _ = ens.get()
        ek = 0.5 * float((v.double() ** 2).sum())
        return ek, ek + 0.5 * float(U.double().sum()) - 0.5 * u_cut * float(nn.double().sum())

    ek0, e0 = shifted_energy()  # also a_0 (and the first sizing rebuild)
    F0, nn0 = ens.sorted_results(L.OFP_PREC_F32C)
    f0 = torch.linalg.vector_norm(F0, dim=0)
    st0 = ens.status()
    steps = args.dyn_steps
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    ens.run(L.OFP_PREC_F32C, wl.params["dt"], steps)
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = s.elapsed_time(e)
    st = ens.status()
    p, v, F = ens.get()
    speed = torch.linalg.vector_norm(v.double(), dim=0)
    outside = ((p < wl.box_lo[0]) | (p >= wl.box_hi[0])).any(dim=0)
    _, e1 = shifted_energy()
    t_rebuild = _evt_ms(lambda: ens.rebuild(sizing=False), 2)
    out = {"workload": "power-law clusters, LJ velocity Verlet (LJ units)", "n_particles": wl.params["n"],
           "box_side": wl.params["box"], "r_cut": wl.r_cut, "sigma": wl.params["sigma"], "dt": wl.params["dt"],
           "init_min_distance": wl.params["d_min"], "init_redraw_rounds": rounds, "init_pairs_below_min": n_close, "init_s": t_init,
           "steps": steps, "ms_per_step": ms / steps,
           "wall_s": wall, "rebuilds": int(st["rebuilds"] - st0["rebuilds"]), "n_ghost": int(st["n_ghost"]),
           "verlet_entries": int(st["n_entries"]), "pairs_initial": int(nn0.sum().item()),
           "initial_force_max": float(f0.max()), "initial_force_median": float(f0.median()),
           "initial_nonfinite_forces": int((~torch.isfinite(F0)).any(dim=0).sum().item()),
           "final_nonfinite_positions": int((~torch.isfinite(p)).any(dim=0).sum().item()),
           "final_speed_max": float(speed[torch.isfinite(speed)].max()),
           "final_particles_speed_gt_1e3": int((speed > 1e3).sum().item()),
           "final_outside_box_until_next_rebuild": int(outside.sum().item()),
           "energy_shifted": {"initial": e0, "final": e1, "kinetic_initial": ek0,
                              "drift_over_kinetic": abs(e1 - e0) / ek0},
           "rebuild_ms": t_rebuild,
           "recaptures": int(st["recaptures"])}
    del ens, p, v, F, F0
    torch.cuda.empty_cache()
    return out


def ensemble_section(args):
    """Batching study: 256 periodic LJ boxes (log-uniform sizes in [1024, 262144],
    r_c ~ U[0.03, 0.08], sigma = 0.4 r_c, uniform FP32).  The FP32-compensated
    batched form (one launch per pass for all boxes, autotuned) beside the plain
    FP64 form (one vector_dist and one LJ launch per box) on the same boxes."""
    import torch

    from openfpm_pdata_amd import _lib as L
    from openfpm_pdata_amd import autotune, workloads
    from openfpm_pdata_amd import ensemble as E
    from openfpm_pdata_amd import vector_dist as vdm

    wls = workloads.ensemble(n_boxes=args.ens_boxes)
    systems = [E.System(wl.n, wl.r_cut, wl.params["sigma"], 1.0, 1.0, tuple(wl.box_lo), tuple(wl.box_hi))
               for wl in wls]
    pos = torch.from_numpy(np.ascontiguousarray(np.concatenate([wl.pos for wl in wls], axis=1))).cuda()
    n_tot = int(pos.shape[1])
    dt, vmax = 1e-4, 5.0  # Maxwell T = 1, m = 1: fastest of ~1e7 particles ~ 5 sigma_v
    opts, log = autotune.tune(systems, pos, L.OFP_PREC_F32C, dt, v_max=vmax, budget_s=args.tune_budget)
    torch.cuda.synchronize()
    out = {"workload": "ensemble of periodic LJ boxes", "n_boxes": len(wls), "n_particles": n_tot,
           "n_min": int(min(s.n for s in systems)), "n_max": int(max(s.n for s in systems)),
           "autotune": {"chosen": dict(vars(opts)), "objective": "force pass + rebuild / steps between rebuilds "
                        f"(v_max {vmax}, dt {dt})", "candidates": log}}
    ens = E.Ensemble(systems, torch.float32, opts)
    ens.set(pos)
    ens.rebuild()
    res = {}
    for prec, name in ((L.OFP_PREC_F32C, "fp32c_batched"), (L.OFP_PREC_F64, "fp64_batched")):
        t_forces = _evt_ms(lambda: ens.force_pass(prec), 3)
        t_pair = _evt_ms(lambda: ens.pair_pass(prec), 5)
        F, nn = ens.sorted_results(prec)
        res[prec] = (F, nn)
        # one hipGraph replay of the full step (kick-drift, ghost refresh, reorder,
        # pair pass, kick); dt = 0 keeps the configuration, so no rebuild is due
        ens.run(prec, 0.0, 2)
        t_step = _evt_ms(lambda: ens.run(prec, 0.0, 1), 5)
        t_reb = _evt_ms(lambda: ens.rebuild(sizing=False), 2)
        out[name] = {"pair_kernel_ms": t_pair, "force_pass_ms": t_forces, "step_graph_ms": t_step,
                     "rebuild_ms": t_reb}
    pairs = int(res[L.OFP_PREC_F64][1].sum().item())
    nonfinite = int((~torch.isfinite(res[L.OFP_PREC_F32C][0])).sum().item())
    F32, F64 = res[L.OFP_PREC_F32C][0], res[L.OFP_PREC_F64][0]
    d = torch.linalg.vector_norm(F32 - F64, dim=0)
    mag = torch.linalg.vector_norm(F64, dim=0)
    e = d / torch.clamp(mag, min=float(torch.sqrt((mag ** 2).mean())))
    out["pairs"] = pairs
    out["nn_mismatch"] = int((res[L.OFP_PREC_F32C][1] != res[L.OFP_PREC_F64][1]).sum().item())
    out["error_vs_fp64"] = {"force_max": float(e.max()), "force_rms": float(torch.sqrt((e ** 2).mean())),
                            "nonfinite_fp32c": nonfinite,
                            "reference": "FP64 arithmetic on the same positions, ghosts and Verlet lists"}
    nb = autotune.steps_between(systems, np.full(len(systems), vmax), opts.skin, dt)
    for name in ("fp32c_batched", "fp64_batched"):
        r = out[name]
        r["pairs_per_s"] = pairs / (r["pair_kernel_ms"] * 1e-3)
        r["step_amortised_ms"] = r["step_graph_ms"] + r["rebuild_ms"] / nb
        r["steps_between_rebuilds"] = nb
    st = ens.status()
    out["verlet_entries"] = int(st["n_entries"])
    out["n_ghost"] = int(st["n_ghost"])
    del ens, F32, F64, res
    torch.cuda.empty_cache()

    # plain FP64: one vector_dist per box, lists as the batched form, one LJ launch per box
    rvf = 1.0 + opts.skin
    boxes = []
    for wl in wls:
        rv = wl.r_cut * rvf
        vd = vdm.VectorDistGPU(wl.n, wl.box_lo, wl.box_hi, wl.bc, rv, dtype=torch.float32,
                               capacity=int(wl.n * 2.2) + 1024)
        vd.pos[:, : wl.n] = torch.from_numpy(wl.pos).cuda()
        vd.ghost_get()
        cl, vl = vd.getVerletGPU(rv, vdm.CL_NON_SYMMETRIC, sub=opts.sub)
        F = torch.empty((3, vd.cap), dtype=torch.float64, device="cuda")
        prm = L.LJParams(wl.r_cut, wl.params["sigma"], 1.0)
        boxes.append((vd, cl, vl, F, prm))
    st_ = vdm._stream()

    def plain():
        for vd, cl, vl, F, prm in boxes:
            fp = vdm.planes_of(F)
            L.call("ofp_pair_lj", cl.h, vl.h, vd.pos.data_ptr(), vd.cap, L.OFP_PREC_F64, prm, fp[0][0], fp[1][0],
                   fp[2][0], None, None, st_)

    t_plain = _evt_ms(plain, 3)
    out["fp64_plain"] = {"pair_kernels_ms": t_plain, "launches": len(boxes), "pairs_per_s": pairs / (t_plain * 1e-3)}
    del boxes
    torch.cuda.empty_cache()
    out["speedup_fp32c_batched_over_fp64_plain"] = out["fp32c_batched"]["pairs_per_s"] / out["fp64_plain"]["pairs_per_s"]
    out["speedup_fp32c_batched_over_fp64_batched"] = (out["fp32c_batched"]["pairs_per_s"] /
                                                     out["fp64_batched"]["pairs_per_s"])
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from openfpm_pdata_amd import _lib as L
    from openfpm_pdata_amd import workloads

    wl = workloads.clustered_sph(n=args.n)
    step = Step(wl, L.OFP_PREC_F32C, args.sub)
    ms, wall = timed_steps(step, args.warmup, args.steps)
    phases = step.phase_ms()
    pairs, G, E = step.n_pairs()
    N_tot = wl.n + G
    value = 2 * pairs / (ms * 1e-3)
    err = error_vs_fp64(step)

    res = {
        "metric": "pair interactions/s + CellList build Mpart/s, FP32-comp vs FP64, 1 GPU",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "clustered_sph", "n_particles": wl.n, "n_ghost": G, "h": wl.params["h"],
                   "r_cut": wl.r_cut, "verlet_skin": wl.params["skin"], "pairs_per_pass": pairs,
                   "verlet_entries": E, "precision": "fp32-compensated"},
        "celllist_build_mpart_s": N_tot / (phases["cell_list"] * 1e-3) / 1e6,
        "phases_ms": phases,
        "wall_s": wall,
        "error_vs_fp64": err,
    }
    # dominant kernel roofline: the SPH force pass over the Verlet list
    force_bytes = E * 4 + (wl.n + 1) * 8 + N_tot * (16 + 4) + wl.n * 12
    achieved = force_bytes / (phases["force"] * 1e-3) / 1e9
    res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "sph_force (Verlet)",
                       "algorithmic_bytes": force_bytes}
    pmc = load_traffic("ofp::k_pairs_f32c<ofp::Force32, true, true>")  # Verlet rows, cubic cells
    if pmc:
        res["roofline"]["traffic"] = pmc["bytes"]
        res["roofline"]["traffic_source"] = pmc["source"]
    # the step's cell list (cells of r_c, positions reordered, quantised stream):
    # positions read and written (3s each), sort_to_nonsort, nonsort_to_sort and
    # the quantised record (4 + 4 + 16 B) per particle, domain ids (4 B per domain
    # particle), cell starts (4 B per cell + 1)
    C_step = int(step.cl.view().n_cells)
    cl_bytes = N_tot * (12 + 12 + 4 + 4 + 16) + wl.n * 4 + 4 * (C_step + 1)
    cl_gbs = cl_bytes / (phases["cell_list"] * 1e-3) / 1e9
    res["roofline_celllist"] = {"bound": "hbm", "achieved": cl_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": cl_gbs / HBM_PEAK_GBS, "algorithmic_bytes": cl_bytes,
                                "n_particles": N_tot, "n_cells": C_step}
    del step
    torch.cuda.empty_cache()

    if not args.no_fp64:
        step64 = Step(wl, L.OFP_PREC_F64, args.sub)
        ms64, _ = timed_steps(step64, 1, max(1, args.steps // 2))
        ph64 = step64.phase_ms()
        res["fp64"] = {"value": 2 * pairs / (ms64 * 1e-3), "ms_per_step": ms64, "phases_ms": ph64,
                       "speedup_fp32c_over_fp64": ms64 / ms}
        del step64
        torch.cuda.empty_cache()

    def section(name, fn):
        try:
            res[name] = fn(args)
        except Exception as ex:  # a failing study is reported, the headline line still prints
            res[name] = {"error": f"{type(ex).__name__}: {ex}"}
            torch.cuda.synchronize()
            torch.cuda.empty_cache()

    if args.cl_scale_n > 0:
        section("celllist_at_scale", celllist_at_scale)
    if not args.no_remote:
        section("ghost_remote", remote_section)
    if not args.no_map:
        section("map", map_section)
    if not args.no_sym:
        section("symmetric", sym_section)
    if not args.no_sparse:
        section("sparse_cell_list", sparse_section)
    if not args.no_ensemble:
        section("ensemble", ensemble_section)
    if args.dyn_n > 0:
        section("dynamic", dynamic_section)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(wl, min(args.cpu_sample, wl.n))
    if args.copy_probe:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        res["copy_probe_bytes"] = a.numel() * 4
    print(json.dumps(res))


if __name__ == "__main__":
    main()
