#!/usr/bin/env python3
"""Benchmark of the vector_dist neighbour-search hot path on one MI355X.

Headline (the dynamic run, the largest configuration that runs on one GPU):
67,108,864 FP32 particles from 512 power-law-weighted Gaussian clusters in a
periodic box, Lennard-Jones in LJ units (sigma = eps = m = 1, r_c = 2.5),
Maxwell velocities at T = 1, velocity Verlet at dt = 1e-4.  One step = kick and
drift, ghost refresh (ghost_get with SKIP_LABELLING), reorder into cell order
with the quantised stream, LJ over the Verlet rows (FP32 storage, pairs
evaluated in FP64), kick; every step is one hipGraph replay.  The Verlet list
is rebuilt (map, ghost_get, updateCellListGPU, Verlet build) when a particle
may have moved beyond skin r_c / 2; skin, cells and tile shape come from the
autotuner.  The timed region is a run of K steps from the run's own neighbour
build: the initial rebuild plus K steps with every rebuild the displacement
trigger fires.  `value` = pair interactions per second over that region
(ordered pairs i != j with d < r_c, per step, times K, over its time).
The same run in the plain FP64 form is reported beside it, with the FP32-storage
form's force error against FP64.

Sections: the clustered SPH step (4,194,304 particles), the cell list alone on
the dynamic-run particles, the batching study (256 boxes), the one-rank send
sides of ghost_get and map, the symmetric list, the sparse cell list.  The CPU
baseline is the oracle (C, one thread, -O3 -march=native on this host).

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


def progress(msg: str) -> None:
    """Progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed velocity-Verlet steps of the dynamic run")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4194304, help="particles of the clustered SPH section")
    ap.add_argument("--sph-steps", type=int, default=10, help="timed steps of the clustered SPH section")
    ap.add_argument("--no-sph", action="store_true", help="skip the clustered SPH section")
    ap.add_argument("--cpu-runs", type=int, default=30, help="timed runs of each CPU baseline measurement")
    ap.add_argument("--dyn-tune-budget", type=float, default=30.0, help="autotuner budget of the dynamic run (s)")
    ap.add_argument("--dyn-opts", type=str, default=None,
                    help="ensemble options of the dynamic run as JSON (skips its autotuner; for profiling runs)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sub", type=int, default=2, help="Verlet cell list: cells of r/sub, stencil sub")
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with a 1 GiB device-to-device copy (calibrates HBM byte counters)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batching study")
    ap.add_argument("--no-remote", action="store_true", help="skip the remote-ghost (send side) study")
    ap.add_argument("--no-map", action="store_true", help="skip the map (migration) study")
    ap.add_argument("--no-sym", action="store_true", help="skip the symmetric-list study")
    ap.add_argument("--no-sparse", action="store_true", help="skip the sparse cell-list study")
    ap.add_argument("--dyn-n", type=int, default=1 << 26, help="particles of the dynamic run")
    ap.add_argument("--ens-boxes", type=int, default=256)
    ap.add_argument("--tune-budget", type=float, default=40.0, help="autotuner time budget (s)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT_JSON"), default=None,
                    help="no benchmark: summarise two rocprofv3 --pmc runs of this benchmark (FETCH_SIZE, then "
                         "WRITE_SIZE, each with --copy-probe) into per-kernel HBM bytes per dispatch")
    return ap.parse_args()


class Step:
    """One pass of the hot path on one vector (FP32C or FP64)."""

    def __init__(self, wl, precision, sub=2, walk="verlet"):
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
        # the step's cell list: cells of r_verlet / sub, stencil sub (the Verlet
        # list's, getVerlet's geometry)
        self.clv = self.vd.getCellListGPU(self.r_verlet / sub, opt, NNIteratorBox=sub,
                                          ghostEnlargeFactor=vdm.GHOST_ENLARGE_VL)
        self.vl = vdm.VerletListGPU()
        # "verlet": the pair passes read the Verlet list built each step; "tiles":
        # no Verlet list, the passes walk the cell list in LDS-staged tiles
        self.walk = walk
        self.ev = {}
        self.times = {}

    def _mark(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.ev.setdefault(name, []).append(e)

    def run(self, record=False):
        vd, L, vdm, clv = self.vd, self.L, self.vdm, self.clv
        st = vdm._stream()
        m = self._mark if record else (lambda name: None)
        m("t0")
        vd.ghost_get()
        m("ghost")
        vd.updateCellListGPU(clv)
        m("cell_list")
        vlh = None
        if self.walk == "verlet":
            self.vl.build(clv, vd.pos, self.r_verlet)
            vlh = self.vl.h
        m("verlet")
        rho, prs, pterm, F = (vd.prp[k] for k in ("rho", "prs", "pterm", "force"))
        L.call("ofp_sph_density", clv.h, vlh, vd.pos.data_ptr(), vd.cap, self.prec, self.prm, rho.data_ptr(), st)
        m("density")
        vd.ghost_sync_sorted(clv, ("rho",))
        L.call("ofp_sph_eos", self.prec, rho.data_ptr(), vd.size, self.prm, prs.data_ptr(), pterm.data_ptr(), st)
        m("eos")
        fp = vdm.planes_of(F)
        L.call("ofp_sph_force", clv.h, vlh, vd.pos.data_ptr(), vd.cap, self.prec, self.prm, pterm.data_ptr(),
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
    vlh = None
    if step.walk == "verlet":
        step.vl.build(clv, vd.pos, step.r_verlet)
        vlh = step.vl.h
    out = {}
    for prec, pdt in ((L.OFP_PREC_F32C, torch.float32), (L.OFP_PREC_F64, torch.float64)):
        rho = torch.zeros(vd.cap, dtype=torch.float64, device="cuda")
        pterm = torch.zeros(vd.cap, dtype=pdt, device="cuda")
        F = torch.zeros((3, vd.cap), dtype=pdt, device="cuda")
        fp = vdm.planes_of(F)
        L.call("ofp_sph_density", clv.h, vlh, vd.pos.data_ptr(), vd.cap, prec, step.prm, rho.data_ptr(), st)
        L.call("ofp_ghost_sync_sorted", vd._gh, clv.h, vdm.plane_array([(rho.data_ptr(), 8)]), 1, st)
        L.call("ofp_sph_eos", prec, rho.data_ptr(), vd.size, step.prm, None, pterm.data_ptr(), st)
        L.call("ofp_sph_force", clv.h, vlh, vd.pos.data_ptr(), vd.cap, prec, step.prm, pterm.data_ptr(),
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
            "reference": "FP64 arithmetic on the same positions, ghosts and lists"}


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


def pmc_summary(fetch_dir: str, write_dir: str, out_json: str) -> None:
    """Per-kernel FETCH_SIZE and WRITE_SIZE (KiB per dispatch, the median over
    the kernel's dispatches: the first steps of a run that starts from the
    caller's order, before any rebuild has put the particles in cell order,
    are not typical) from the counter_collection CSVs of two rocprofv3 --pmc
    runs, with the factors that turn each into bytes: 1 GiB over what the
    counter reads for the 1 GiB device copy of --copy-probe (the largest
    copyBuffer dispatch).  Kernel names without the argument list."""
    import collections
    import csv
    import glob
    import statistics

    def load(d, counter):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        per = {}
        for f in files:
            for r in csv.DictReader(open(f)):
                if r.get("Counter_Name") != counter:
                    continue
                key = (f, r.get("Dispatch_Id") or r.get("Correlation_Id"))
                name = r["Kernel_Name"].split("(")[0]
                name = name[5:] if name.startswith("void ") else name
                prev = per.get(key, (name, 0.0))[1]
                per[key] = (name, prev + float(r["Counter_Value"]))
        agg = collections.defaultdict(list)
        for name, v in per.values():
            agg[name].append(v)
        copies = [max(v) for k, v in agg.items() if "copyBuffer" in k]
        if not copies:
            raise SystemExit(f"{d}: no copyBuffer dispatch (run with --copy-probe)")
        return {k: (len(v), statistics.median(v)) for k, v in agg.items()}, max(copies)

    fetch, fcopy = load(fetch_dir, "FETCH_SIZE")
    write, wcopy = load(write_dir, "WRITE_SIZE")
    out = {"fetch_factor": 1048576.0 / fcopy, "write_factor": 1048576.0 / wcopy, "copy_fetch_kib": fcopy,
           "copy_write_kib": wcopy, "kernels": {}}
    for k in fetch:
        if k in write and "copyBuffer" not in k:
            out["kernels"][k] = {"dispatches": fetch[k][0], "fetch_kib": fetch[k][1], "write_kib": write[k][1]}
    json.dump(out, open(out_json, "w"), indent=1)
    ff, wf = out["fetch_factor"], out["write_factor"]
    rows = sorted(out["kernels"].items(), key=lambda kv: -(kv[1]["fetch_kib"] * ff + kv[1]["write_kib"] * wf))
    print(f"factors: fetch {ff:.4f}, write {wf:.4f}")
    print("| kernel | dispatches | read GB / dispatch | written GB / dispatch |")
    print("|---|---|---|---|")
    for k, v in rows[:25]:
        print(f"| `{k}` | {v['dispatches']} | {v['fetch_kib'] * ff * 1024 / 1e9:.3f} | "
              f"{v['write_kib'] * wf * 1024 / 1e9:.3f} |")


def steady_step_ms(ens, prec) -> float:
    """One step of a run in the steady state (a graph replay that is neither
    the run's first step nor its last): runs of 8 and of 2 steps differ by six
    such steps.  dt = 0 keeps the configuration, so no rebuild is due."""
    t2 = _evt_ms(lambda: ens.run(prec, 0.0, 2), 3)
    t8 = _evt_ms(lambda: ens.run(prec, 0.0, 8), 3)
    return max(t8 - t2, 0.0) / 6


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


def _celllist_builds(pos, n_dom, blo, bhi, div, r):
    """Two builds of one cell list: from the particles' given order, then from
    the cell order the first build produced (the order a run that keeps its
    particles sorted feeds in).  Algorithmic bytes per build: the positions read
    and written in cell order (3s each), sort_to_nonsort and nonsort_to_sort
    (4 B each) per particle, the domain sort ids (4 B per domain particle) and
    the cell starts (4 B per cell + 1): N·(6s + 8) + 4·N_dom + 4(C+1)."""
    import torch

    from openfpm_pdata_amd import vector_dist as vdm

    n = int(pos.shape[1])
    cl = vdm.CellListGPU(torch.float32, blo, bhi, div, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION, r)
    out_pos = torch.empty_like(pos)
    C = int(np.prod(np.asarray(div, np.int64) + 2))
    alg = n * (12 + 12 + 8) + 4 * n_dom + 4 * (C + 1)
    res = {"n_particles": n, "n_cells": C, "algorithmic_bytes": alg}
    for name in ("random_order", "cell_order"):
        t = _evt_ms(lambda: cl.construct(pos, n, n_dom, out_pos), 3)
        res[name] = {"ms": t, "mpart_s": n / (t * 1e-3) / 1e6, "achieved_gbs": alg / (t * 1e-3) / 1e9,
                     "frac_hbm": alg / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
        pos = out_pos.clone()  # next build reads the sorted positions
    del cl, pos, out_pos
    torch.cuda.empty_cache()
    return res


def celllist_at_scale(args):
    """updateCellListGPU on 2^26 uniform FP32 particles in the periodic unit cube
    (~9 per cell, 195^3 cells): a working set well beyond the 256 MB Infinity
    Cache; from random order, then from cell order (see _celllist_builds)."""
    import torch

    from openfpm_pdata_amd import vector_dist as vdm

    n = 1 << 26
    g = torch.Generator(device="cuda").manual_seed(0x0F9D0005)
    pos = torch.rand((3, n), generator=g, device="cuda", dtype=torch.float32)
    r = (9.0 / n) ** (1.0 / 3.0)
    blo, bhi, div = vdm.cl_param(torch.float32, (0.0,) * 3, (1.0,) * 3, r, [0.0] * 3, [0.0] * 3, 1.0)
    res = {"workload": "2^26 uniform FP32 particles, ~9 per cell"}
    res.update(_celllist_builds(pos, n, blo, bhi, div, r))
    return res


def celllist_dynamic(args, ctx):
    """updateCellListGPU on the dynamic run's particles (2^26 domain particles
    in 512 clusters and their ghosts, FP32; cells of r_c over the box enlarged
    by the ghost, most of them empty): from the generator's order (random with
    respect to the cells), then from cell order (see _celllist_builds)."""
    import torch

    from openfpm_pdata_amd import vector_dist as vdm

    pos = ctx["P"].contiguous()
    r, g = ctx["r_cut"], ctx["ghost"]
    blo, bhi, div = vdm.cl_param(torch.float32, ctx["box_lo"], ctx["box_hi"], r, [-g] * 3, [g] * 3,
                                 vdm.GHOST_ENLARGE_CL)
    res = {"workload": "dynamic-run particles (domain + ghosts), cells of r_c"}
    res.update(_celllist_builds(pos, ctx["N"], blo, bhi, div, r))
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


DYN_SKINS = (0.01, 0.02, 0.05, 0.1)


def _host_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def _mean_std(xs):
    a = np.asarray(xs, np.float64)
    return float(a.mean()), float(a.std(ddof=1)) if len(a) > 1 else 0.0


def dynamic_headline(args):
    """The dynamic run (see the module docstring).  Returns (fields of the
    headline line, context kept for the cell-list section and the CPU
    baseline)."""
    import torch

    from openfpm_pdata_amd import _lib as L
    from openfpm_pdata_amd import autotune
    from openfpm_pdata_amd import ensemble as E
    from openfpm_pdata_amd import workloads

    t_init = time.perf_counter()
    wl, pos, vel, rounds, n_close = workloads.power_law_clusters_lj(n=args.dyn_n)
    t_init = time.perf_counter() - t_init
    dt = wl.params["dt"]
    sy = E.System(wl.params["n"], wl.r_cut, wl.params["sigma"], 1.0, 1.0, tuple(wl.box_lo), tuple(wl.box_hi))
    vmax = float(torch.linalg.vector_norm(vel.double(), dim=0).max())
    # autotuned options: the rebuild period follows the fastest particle; the
    # run's skins are finer than the generic set (dt is small against the cell)
    progress(f"dynamic run: workload ready ({t_init:.1f} s), autotuning")
    t_tune = time.perf_counter()
    if args.dyn_opts:
        opts, log = E.Options(**json.loads(args.dyn_opts)), "options given"
    else:
        saved = autotune.SKINS
        autotune.SKINS = DYN_SKINS
        try:
            opts, log = autotune.tune([sy], pos, L.OFP_PREC_F32C, dt, v_max=vmax,
                                      start=E.Options(skin=0.05, sub=1, lanes=16, waves_per_cu=32),
                                      budget_s=args.dyn_tune_budget)
        finally:
            autotune.SKINS = saved
    t_tune = time.perf_counter() - t_tune
    torch.cuda.empty_cache()
    ens = E.Ensemble([sy], torch.float32, opts)
    ens.set(pos, vel)
    u_cut = 4.0 * (0.4 ** 12 - 0.4 ** 6)  # u(r_c) at r_c = 2.5 sigma

    def shifted_energy(prec):
        _, U = ens.forces(prec, energy=True)
        _, nn = ens.sorted_results(prec)
        _, v, _ = ens.get()
        ek = 0.5 * float((v.double() ** 2).sum())
        return ek, ek + 0.5 * float(U.double().sum()) - 0.5 * u_cut * float(nn.double().sum()), nn

    ek0, e0, nn0 = shifted_energy(L.OFP_PREC_F32C)  # sizing rebuild, a_0
    pairs = int(nn0.sum().item())
    st0 = ens.status()
    N, G, E_ = int(st0["n_dom"]), int(st0["n_ghost"]), int(st0["n_entries"])
    # storage-order positions (domain + ghosts) for the cell-list section and the CPU baseline
    a = ens.arrays()
    from openfpm_pdata_amd.vector_dist import device_view

    P_dev = device_view(a.pos, 3 * int(a.cap), torch.float32).view(3, -1)[:, : N + G].clone()
    _, divc, lo, sp = ens.cells()

    def timed_run(prec, steps, warmup):
        ens.run(prec, dt, warmup)  # graphs captured, caches warm
        torch.cuda.synchronize()
        r0 = ens.status()["rebuilds"]
        s_, e_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        s_.record()
        ens.rebuild(sizing=False)  # the run's own neighbour build
        ens.run(prec, dt, steps)
        e_.record()
        torch.cuda.synchronize()
        return s_.elapsed_time(e_), time.perf_counter() - t0, int(ens.status()["rebuilds"] - r0)

    K, W = args.steps, args.warmup
    progress(f"dynamic run: tuned in {t_tune:.1f} s, timing {K} steps")
    ms, wall, rebuilds = timed_run(L.OFP_PREC_F32C, K, W)
    # the dominant kernel alone: LJ over the Verlet rows (forces only, as a step
    # runs it), HIP events on the library's stream (torch's current stream)
    ens.forces(L.OFP_PREC_F32C)
    t_lj = _evt_ms(lambda: ens.pair_pass(L.OFP_PREC_F32C), 5)
    t_rebuild = _evt_ms(lambda: ens.rebuild(sizing=False), 3)
    t_step = steady_step_ms(ens, L.OFP_PREC_F32C)
    st1 = ens.status()
    E1 = int(st1["n_entries"])
    # FP32-storage forces against FP64 arithmetic on the same configuration and lists
    F32, nn32 = None, None
    ens.forces(L.OFP_PREC_F32C, energy=True)
    F32, nn32 = ens.sorted_results(L.OFP_PREC_F32C)
    ens.forces(L.OFP_PREC_F64, energy=True)
    F64, nn64 = ens.sorted_results(L.OFP_PREC_F64)
    d = torch.linalg.vector_norm(F32 - F64, dim=0)
    mag = torch.linalg.vector_norm(F64, dim=0)
    err = d / torch.clamp(mag, min=float(torch.sqrt((mag ** 2).mean())))
    nn_mismatch = int((nn32 != nn64).sum().item())
    err = {"force_max": float(err.max()), "force_rms": float(torch.sqrt((err ** 2).mean())),
           "nn_mismatch": nn_mismatch, "reference": "FP64 arithmetic on the same positions, ghosts and lists"}
    del F32, F64, nn32, nn64, d, mag
    # the same run in the plain FP64 form
    progress("dynamic run: the FP64 form")
    ms64, _, reb64 = timed_run(L.OFP_PREC_F64, K, W)
    ens.forces(L.OFP_PREC_F64)
    t_lj64 = _evt_ms(lambda: ens.pair_pass(L.OFP_PREC_F64), 3)
    _, e1, _ = shifted_energy(L.OFP_PREC_F32C)
    if opts.pair_walk == 1:
        # cell tiles: each particle's 16-byte quantised record read once
        # (16 (N + G)), the domain ids of the targets (4 N), the FP32 forces
        # written (12 N)
        lj_bytes = 16 * (N + G) + 4 * N + 12 * N
        kname = "ofp::k_tile_f32c<float, ofp::LJ32T<false>, true, true>"
        kdesc = "LJ over LDS-staged cell tiles of the last rebuild's cells"
    else:
        # Verlet rows: the row stream (4 E), offsets and domain ids of the rows
        # (8 (N + 1) + 4 N), the system id of each row (4 N), each particle's
        # 16-byte quantised record read once (16 (N + G)), the FP32 forces
        # written (12 N)
        lj_bytes = 4 * E1 + 8 * (N + 1) + 8 * N + 16 * (N + G) + 12 * N
        kname = f"ofp::k_pairs_many_f32c<float, ofp::LJ32T<false>, {opts.lanes}, true, ofp::NoKick>"
        kdesc = ("LJ over the Verlet rows, forces stored (a run's steps launch the same kernel with the "
                 "velocity-Verlet kick in place of the store: ofp::StepKick<float>)")
    achieved = lj_bytes / (t_lj * 1e-3) / 1e9
    res = {
        "value": pairs * K / (ms * 1e-3),
        "ms_per_step": ms / K,
        "config": {"workload": "dynamic_run_power_law_clusters_lj", "n_particles": N, "n_ghost": G,
                   "n_clusters": 512, "box_side_sigma": wl.params["box"], "r_cut_sigma": wl.r_cut, "dt": dt,
                   "steps_timed": K, "pairs_per_step": pairs, "verlet_entries": E_,
                   "autotuned": dict(vars(opts)), "precision": "FP32 storage, FP64 pair evaluation",
                   "timed_region": "the run's neighbour build + K velocity-Verlet steps (hipGraph replays) with "
                                   "every rebuild the displacement trigger fires"},
        "dynamic": {"rebuilds_in_timed_region": rebuilds + 1, "triggered_rebuilds": rebuilds,
                    "rebuild_ms": t_rebuild, "step_ms_without_rebuild": t_step, "lj_kernel_ms": t_lj,
                    "lj_pairs_per_s": pairs / (t_lj * 1e-3), "lj_entries_per_s": E1 / (t_lj * 1e-3),
                    "wall_s": wall, "v_max_initial": vmax,
                    "steps_between_rebuilds_predicted": autotune.steps_between([sy], np.array([vmax]),
                                                                               opts.skin, dt),
                    "energy_shifted": {"initial": e0, "final": e1, "kinetic_initial": ek0,
                                       "drift_over_kinetic": abs(e1 - e0) / ek0},
                    "init": {"seconds": t_init, "redraw_rounds": rounds, "pairs_below_min": n_close,
                             "min_distance": wl.params["d_min"]},
                    "autotune": {"seconds": t_tune, "candidates": log}},
        "fp64": {"value": pairs * K / (ms64 * 1e-3), "ms_per_step": ms64 / K, "rebuilds_in_timed_region": reb64 + 1,
                 "lj_kernel_ms": t_lj64, "speedup_fp32_storage_over_fp64": ms64 / ms,
                 "lj_kernel_speedup": t_lj64 / t_lj},
        "error_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kname, "kernel_role": kdesc, "algorithmic_bytes": lj_bytes, "kernel_ms": t_lj},
    }
    ctx = {"P": P_dev, "N": N, "G": G, "lo": lo[0], "sp": sp[0], "divc": divc[0], "r_cut": wl.r_cut,
           "sigma": wl.params["sigma"], "box_lo": tuple(wl.box_lo), "box_hi": tuple(wl.box_hi),
           "ghost": wl.r_cut * (1.0 + opts.skin)}
    del ens, pos, vel
    torch.cuda.empty_cache()
    return res, ctx


def cpu_baseline_dynamic(ctx, runs=30, targets=200000):
    """The oracle on the host cores, one thread, built -O3 -march=native here:
    CellList::fill (ofo_cl_fill) and the 27-cell LJ loop (ofo_lj_walk) on the
    dynamic run's particles.  The fill is timed on the particles of one octant of
    the box (same cells, ~1/8 of the particles); the pair loop on `targets`
    random domain particles of the whole cell list (built once, untimed).
    Mean and sample stddev over `runs` runs of each."""
    from oracle import oracle as ora

    path = ora.select_native()
    P = np.ascontiguousarray(ctx["P"].cpu().numpy())
    N = ctx["N"]

    class Grid:
        def __init__(self, lo, sp, divc):
            self.lo = np.ascontiguousarray(lo, np.float32)
            self.sp = np.ascontiguousarray(sp, np.float32)
            self.divc = np.ascontiguousarray(divc, np.uint32)
            self.n_cells = int(np.prod(self.divc.astype(np.int64)))

    grid = Grid(ctx["lo"], ctx["sp"], ctx["divc"])
    cid = ora.cell_ids(P, grid).astype(np.int64)
    d0, d1 = int(grid.divc[0]), int(grid.divc[1])
    cc = np.stack([cid % d0, (cid // d0) % d1, cid // (d0 * d1)])
    half = (grid.divc.astype(np.int64) - 2) // 2
    octant = np.all((cc >= 1) & (cc <= half[:, None]), axis=0)
    Po = np.ascontiguousarray(P[:, octant])
    go = Grid(ctx["lo"], ctx["sp"], half + 2)
    t_fill = []
    for _ in range(runs):
        t0 = time.perf_counter()
        ora.cl_fill(Po, go)
        t_fill.append(time.perf_counter() - t0)
    cs, ci = ora.cl_fill(P, grid)
    tg = np.random.default_rng(1).choice(N, min(targets, N), replace=False)
    t_pair = []
    pairs = 0
    for _ in range(runs):
        t0 = time.perf_counter()
        _, _, nn = ora.lj_walk(P, tg, grid, cs, ci, ctx["r_cut"], ctx["sigma"], 1.0)
        t_pair.append(time.perf_counter() - t0)
        pairs = int(nn.sum())
    mf, sf = _mean_std(t_fill)
    mp, spair = _mean_std(t_pair)
    n_oct = Po.shape[1]
    out = {"value": pairs / mp, "unit": "pair interactions/s", "cores": 1, "kind": "port",
           "sample": f"oracle (C, one thread, -O3 -march=native) on the dynamic run's particles: LJ over the 27 "
                     f"cells for {len(tg)} random targets ({pairs} pairs), CellList::fill of the {n_oct} "
                     f"particles of one octant; {runs} runs each",
           "pairs_per_s_mean": pairs / mp, "pair_seconds_mean": mp, "pair_seconds_std": spair,
           "celllist_mpart_s": n_oct / mf / 1e6, "fill_seconds_mean": mf, "fill_seconds_std": sf,
           "runs": runs, "build": os.path.relpath(path, ROOT)}
    out.update(_host_info())
    return out


def ensemble_section(args):
    """Batching study: 256 periodic LJ boxes (log-uniform sizes in [1024, 262144],
    r_c ~ U[0.03, 0.08], sigma = 0.4 r_c, uniform FP32).  The FP32-storage
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
    # a second rebuild stores the particles in cell order, as a run's rebuilds keep them
    ens.rebuild(sizing=False)
    res = {}
    for prec, name in ((L.OFP_PREC_F32C, "fp32c_batched"), (L.OFP_PREC_F64, "fp64_batched")):
        t_forces = _evt_ms(lambda: ens.force_pass(prec), 3)
        t_pair = _evt_ms(lambda: ens.pair_pass(prec), 5)
        F, nn = ens.sorted_results(prec)
        res[prec] = (F, nn)
        # one hipGraph replay of a step of a run (ghost slots, the pair pass with
        # the kick); dt = 0 keeps the configuration, so no rebuild is due
        ens.run(prec, 0.0, 2)
        t_step = steady_step_ms(ens, prec)
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


def sph_section(args):
    """The clustered SPH step (4,194,304 FP32 particles, see Step): ghost_get,
    updateCellListGPU (cells of r/2, r = r_c (1 + skin)), Verlet build at r,
    SPH density, sorted-space ghost sync of rho, Tait EOS, SPH force,
    restoreOrder.  `value` = ordered pairs i != j with d < r_c, counted in the
    density and the force pass, over the step time; beside it the same step
    with cell-tile pair passes (no Verlet list) and the FP64 form on an FP64
    copy."""
    import torch

    from openfpm_pdata_amd import _lib as L
    from openfpm_pdata_amd import workloads

    wl = workloads.clustered_sph(n=args.n)
    step = Step(wl, L.OFP_PREC_F32C, args.sub)
    ms, wall = timed_steps(step, 2, args.sph_steps)
    phases = step.phase_ms()
    pairs, G, E = step.n_pairs()
    N_tot = wl.n + G
    res = {"workload": "clustered_sph", "n_particles": wl.n, "n_ghost": G, "h": wl.params["h"],
           "r_cut": wl.r_cut, "verlet_skin": wl.params["skin"], "pairs_per_pass": pairs, "verlet_entries": E,
           "steps": args.sph_steps, "value": 2 * pairs / (ms * 1e-3), "unit": "pair interactions/s",
           "ms_per_step": ms, "phases_ms": phases, "wall_s": wall, "error_vs_fp64": error_vs_fp64(step)}
    # dominant kernel of the step: the SPH force pass over the Verlet list
    force_bytes = E * 4 + (wl.n + 1) * 8 + N_tot * (16 + 4) + wl.n * 12
    achieved = force_bytes / (phases["force"] * 1e-3) / 1e9
    res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "sph_force (Verlet)",
                       "algorithmic_bytes": force_bytes}
    pmc = load_traffic("ofp::k_pairs_f32c<ofp::Force32, true>")
    if pmc:
        res["roofline"]["traffic"] = pmc["bytes"]
        res["roofline"]["traffic_source"] = pmc["source"]
    # the step's Verlet cell list (cells of r/2, reorder, quantised stream)
    C_step = int(step.clv.view().n_cells)
    cl_bytes = N_tot * (12 + 12 + 4 + 4 + 16) + wl.n * 4 + 4 * (C_step + 1)
    cl_gbs = cl_bytes / (phases["cell_list"] * 1e-3) / 1e9
    res["roofline_celllist"] = {"bound": "hbm", "achieved": cl_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": cl_gbs / HBM_PEAK_GBS, "algorithmic_bytes": cl_bytes,
                                "n_particles": N_tot, "n_cells": C_step,
                                "mpart_s": N_tot / (phases["cell_list"] * 1e-3) / 1e6}
    del step
    torch.cuda.empty_cache()
    # the same step with the pair passes over LDS-staged cell tiles of the same
    # cell list (no Verlet list)
    step_t = Step(wl, L.OFP_PREC_F32C, args.sub, walk="tiles")
    ms_t, _ = timed_steps(step_t, 2, args.sph_steps)
    ph_t = step_t.phase_ms()
    res["cell_tiles"] = {"value": 2 * pairs / (ms_t * 1e-3), "ms_per_step": ms_t, "phases_ms": ph_t,
                         "speedup_over_verlet_step": ms / ms_t, "error_vs_fp64": error_vs_fp64(step_t)}
    del step_t
    torch.cuda.empty_cache()
    if not args.no_fp64:
        step64 = Step(wl, L.OFP_PREC_F64, args.sub)
        ms64, _ = timed_steps(step64, 1, max(1, args.sph_steps // 2))
        res["fp64"] = {"value": 2 * pairs / (ms64 * 1e-3), "ms_per_step": ms64, "phases_ms": step64.phase_ms(),
                       "speedup_fp32_storage_over_fp64": ms64 / ms}
        del step64
        torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    progress("dynamic run (headline)")
    head, ctx = dynamic_headline(args)
    progress(f"dynamic run: {head['value']:.4g} pair interactions/s, {head['ms_per_step']:.2f} ms/step")
    res = {
        "metric": "pair interactions/s + CellList build Mpart/s, FP32-comp vs FP64, 1 GPU",
        "value": head["value"],
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage, f64 pair arithmetic",
        "data": "synthetic",
        "config": head["config"],
    }
    res["roofline"] = head["roofline"]
    pmc = load_traffic(head["roofline"]["kernel"])
    if pmc:
        res["roofline"]["traffic"] = pmc["bytes"]
        res["roofline"]["traffic_source"] = pmc["source"]
    for k in ("dynamic", "fp64", "error_vs_fp64"):
        res[k] = head[k]

    def section(name, fn, *a):
        progress(f"section {name}")
        try:
            res[name] = fn(*a)
        except Exception as ex:  # a failing study is reported, the headline line still prints
            res[name] = {"error": f"{type(ex).__name__}: {ex}"}
            torch.cuda.synchronize()
            torch.cuda.empty_cache()

    section("celllist_dynamic", celllist_dynamic, args, ctx)
    ctx["P"] = ctx["P"].cpu()
    torch.cuda.empty_cache()
    section("celllist_at_scale", celllist_at_scale, args)
    cl = res["celllist_dynamic"]
    if "random_order" in cl:
        res["celllist_build_mpart_s"] = {"dynamic_run_random_order": cl["random_order"]["mpart_s"],
                                         "dynamic_run_cell_order": cl["cell_order"]["mpart_s"]}
    if not args.no_sph:
        section("clustered_sph", sph_section, args)
    if not args.no_ensemble:
        section("ensemble", ensemble_section, args)
    if not args.no_remote:
        section("ghost_remote", remote_section, args)
    if not args.no_map:
        section("map", map_section, args)
    if not args.no_sym:
        section("symmetric", sym_section, args)
    if not args.no_sparse:
        section("sparse_cell_list", sparse_section, args)
    if not args.no_cpu:
        progress("cpu baseline")
        res["cpu_baseline"] = cpu_baseline_dynamic(ctx, args.cpu_runs)
    if args.copy_probe:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        res["copy_probe_bytes"] = a.numel() * 4
    print(json.dumps(res))


if __name__ == "__main__":
    main()
