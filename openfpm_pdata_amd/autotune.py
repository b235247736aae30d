"""Autotuner of the ensemble path.

Chooses, by timing them on the workload itself:
  * skin          Verlet radius r_cut (1 + skin): a larger skin lengthens the
                  rows (more pair tests per step) and the time between rebuilds;
  * sub           cells of r_verlet / sub with a stencil of sub cells: smaller
                  cells cut the candidates the Verlet build tests;
  * lanes         tile shape of the pair pass (lanes per Verlet row);
  * waves_per_cu  persistent grid of the Verlet build;
  * verlet_rows   Verlet build kernels: targets in the lanes (dense cells) or
                  one thread per row (dilute cells);
  * pair_walk     Verlet rows, or cell tiles over the last rebuild's cells
                  (no Verlet list: cheaper rebuilds, more candidates per step).
Objective: time per step = one step of a run (a graph replay in the steady
state: after the first step of a run and before its last) + rebuild / (steps
between rebuilds), with the steps between rebuilds bounded by the fastest
particle of each system: floor((skin r_cut / 2) / (v_max dt)).  Coordinate
descent from the defaults, one knob at a time, each candidate measured with
HIP events.
"""
from __future__ import annotations

import math
import time
from dataclasses import asdict, replace

import numpy as np
import torch

from . import _lib as L
from .ensemble import Ensemble, Options, System

SKINS = (0.05, 0.1, 0.2, 0.3)
SUBS = (1, 2, 3)
LANES = (1, 2, 4, 8, 16, 32)
WAVES = (16, 32, 64)
ROWS = (1, 2)  # Verlet build: targets in the lanes, one thread per row
WALKS = (0, 1)  # pair pass: Verlet rows, cell tiles
# knobs that only the Verlet-row walk reads
_VERLET_ONLY = ("lanes", "waves_per_cu", "verlet_rows")


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def steps_between(systems: list[System], vmax: np.ndarray, skin: float, dt: float) -> float:
    """Steps until the fastest particle of some system has moved skin r_cut / 2."""
    n = [0.5 * skin * s.r_cut / max(v * dt, 1e-30) for s, v in zip(systems, vmax)]
    return max(1.0, math.floor(min(n)))


def measure(systems, pos, opts: Options, precision: int, dt: float, vmax, reps: int = 3) -> dict:
    ens = Ensemble(systems, pos.dtype, opts)
    ens.set(pos)
    ens.rebuild()
    # a second rebuild stores the particles in cell order, as a run keeps them
    t_reb = _time(lambda: ens.rebuild(sizing=False), max(1, reps - 1))
    # steady-state steps: runs of 8 and of 2 steps differ by 6 steps that are
    # neither a run's first (records made from the positions) nor its last
    # (forces stored); dt = 0 keeps the configuration, so no rebuild is due
    t2 = _time(lambda: ens.run(precision, 0.0, 2), reps)
    t8 = _time(lambda: ens.run(precision, 0.0, 8), reps)
    t_pass = max(t8 - t2, 0.0) / 6
    nb = steps_between(systems, vmax, opts.skin, dt)
    st = ens.status()
    del ens
    torch.cuda.empty_cache()
    return {"opts": asdict(opts), "pass_ms": t_pass, "rebuild_ms": t_reb, "steps_between": nb,
            "step_ms": t_pass + t_reb / nb, "entries": int(st["n_entries"])}


def tune(systems: list[System], pos: torch.Tensor, precision: int = L.OFP_PREC_F32C, dt: float = 1e-4,
         vel: torch.Tensor | None = None, v_max: float = 5.0, start: Options | None = None,
         budget_s: float = 60.0) -> tuple[Options, list[dict]]:
    """Returns the best options and every measurement.  vel (or v_max for all
    systems) sets the rebuild period."""
    first = np.concatenate([[0], np.cumsum([s.n for s in systems])])
    if vel is not None:
        sp = torch.linalg.vector_norm(vel.double(), dim=0).cpu().numpy()
        vmax = np.array([sp[a:b].max() if b > a else 0.0 for a, b in zip(first[:-1], first[1:])])
    else:
        vmax = np.full(len(systems), v_max)
    best = start or Options()
    log: list[dict] = []
    seen: dict = {}
    t0 = time.perf_counter()

    def cost(o: Options) -> float:
        d = asdict(o)
        if o.pair_walk == 1:  # the cell-tile walk ignores these: one measurement serves all
            for k in _VERLET_ONLY:
                d[k] = None
        key = tuple(d.values())
        if key not in seen:
            try:
                r = measure(systems, pos, o, precision, dt, vmax)
            except L.OfpError as ex:  # a geometry the library rejects (too many cells, ...)
                torch.cuda.empty_cache()
                r = {"opts": asdict(o), "rejected": str(ex), "step_ms": None}
            log.append(r)
            seen[key] = math.inf if r["step_ms"] is None else r["step_ms"]
        return seen[key]

    cur = cost(best)
    for knob, values in (("pair_walk", WALKS), ("verlet_rows", ROWS), ("lanes", LANES), ("sub", SUBS),
                         ("skin", SKINS), ("waves_per_cu", WAVES)):
        for v in values:
            if time.perf_counter() - t0 > budget_s:
                break
            cand = replace(best, **{knob: v})
            c = cost(cand)
            if c < cur:
                best, cur = cand, c
    return best, log
