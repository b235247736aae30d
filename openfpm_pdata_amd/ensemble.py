"""Ensembles of independent periodic LJ systems on one GPU (the ofp_ens_* C ABI).

The reference runs one ``vector_dist`` per system; here the systems share one
set of launches: ghosts of all systems from one count/scan/emit, one cell list
over the grids of all systems (global cell ids, one radix sort), one Verlet
build over a device-side work queue of (system, cell-row) items, one LJ pass,
and velocity Verlet, with the rebuild and the step captured as hipGraphs.  Per
system the ghosts, cell order, Verlet rows and pair arithmetic are those of
the one-system path (``vector_dist.VectorDistGPU``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from .vector_dist import _stream, device_view


@dataclass
class System:
    n: int
    r_cut: float
    sigma: float
    eps: float = 1.0
    mass: float = 1.0
    box_lo: tuple = (0.0, 0.0, 0.0)
    box_hi: tuple = (1.0, 1.0, 1.0)


@dataclass
class Options:
    skin: float = 0.1        # Verlet radius r_cut (1 + skin)
    sub: int = 2             # Verlet cell list: cells of r_verlet / sub, stencil sub
    lanes: int = 16          # lanes per Verlet row in the pair pass (tile shape): 1, 2, 4, 8, 16 or 32
    waves_per_cu: int = 32   # persistent grid of the Verlet build
    headroom: float = 1.25   # capacities over the counts of a sizing rebuild
    verlet_rows: int = 0     # Verlet build: 0 chosen by occupancy, 1 targets in lanes, 2 thread per row
    pair_walk: int = 0       # pair pass: 0 Verlet rows, 1 cell tiles of the last rebuild's cells (no Verlet list)
    split_kick: int = 0      # run steps over Verlet rows: 0 the kick inside the pair kernel, 1 a kernel of its own

    def c(self) -> L.EnsOpts:
        return L.EnsOpts(self.skin, self.sub, self.lanes, self.waves_per_cu, self.verlet_rows, self.headroom,
                         self.pair_walk, self.split_kick)


class Ensemble:
    def __init__(self, systems: list[System], dtype=torch.float32, opts: Options | None = None):
        L.lib()
        self.systems = list(systems)
        self.dtype = dtype
        self.opts = opts or Options()
        self.S = len(systems)
        arr = (L.SysDesc * self.S)()
        for k, s in enumerate(systems):
            arr[k] = L.SysDesc((C.c_double * 3)(*s.box_lo), (C.c_double * 3)(*s.box_hi), s.r_cut, s.sigma, s.eps,
                               s.mass, int(s.n))
        h = C.c_void_p()
        L.call("ofp_ens_create", C.byref(h), L.OFP_F32 if dtype == torch.float32 else L.OFP_F64, self.S, arr,
               C.byref(self.opts.c()))
        self.h = h
        first = (C.c_uint64 * (self.S + 1))()
        L.call("ofp_ens_first", self.h, first)
        self.first = np.array(first[:], dtype=np.int64)
        self.N = int(self.first[-1])

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and getattr(L, "_lib", None) is not None:
            L.lib().ofp_ens_destroy(h)
            self.h = None

    # -------------------------------------------------------------- data
    def set(self, pos: torch.Tensor, vel: torch.Tensor | None = None) -> None:
        """pos, vel: [3, N] device tensors of the ensemble dtype (systems concatenated)."""
        assert pos.shape == (3, self.N) and pos.dtype == self.dtype and pos.is_cuda
        pos = pos.contiguous()
        vel = None if vel is None else vel.contiguous()
        self._keep = (pos, vel)
        L.call("ofp_ens_set", self.h, pos.data_ptr(), None if vel is None else vel.data_ptr(), self.N, _stream())

    def get(self):
        """(pos, vel, force) in the order given to set()."""
        out = [torch.empty((3, self.N), dtype=self.dtype, device="cuda") for _ in range(3)]
        L.call("ofp_ens_get", self.h, *(t.data_ptr() for t in out), self.N, _stream())
        return tuple(out)

    def arrays(self) -> L.EnsArrays:
        a = L.EnsArrays()
        L.call("ofp_ens_arrays_get", self.h, C.byref(a))
        return a

    # ------------------------------------------------------------ passes
    def rebuild(self, sizing: bool = True) -> None:
        L.call("ofp_ens_rebuild", self.h, int(sizing), _stream())

    def force_pass(self, precision: int = L.OFP_PREC_F32C) -> None:
        """Ghost refresh, reorder, LJ, forces back to storage order (no copy out)."""
        L.call("ofp_ens_forces", self.h, precision, None, _stream())

    def forces(self, precision: int = L.OFP_PREC_F32C, energy: bool = False):
        """LJ forces (and per-particle energies) on the current positions, in the
        order given to set()."""
        en = torch.empty(self.N, dtype=self.dtype, device="cuda") if energy else None
        L.call("ofp_ens_forces", self.h, precision, None if en is None else en.data_ptr(), _stream())
        _, _, F = self.get()
        if not energy:
            return F
        gid = self.gid().long()
        e_user = torch.empty_like(en)
        e_user[gid] = en
        return F, e_user

    def pair_pass(self, precision: int, with_energy: bool = False) -> None:
        """The LJ kernel alone (after forces()), for timing: forces only (as a step
        runs it) unless with_energy."""
        L.call("ofp_ens_pair_pass", self.h, precision, int(with_energy), _stream())

    def sorted_results(self, precision: int):
        """(forces [3, N] float64, neighbour counts [N]) of the last pair pass, per
        domain row (rows are the same for both precisions)."""
        a = self.arrays()
        cv, _ = self.lists()
        dom = device_view(cv.dom_sort_ids, self.N, torch.int32).long()
        tdt = torch.float64 if precision == L.OFP_PREC_F64 else torch.float32
        F = device_view(a.force_sorted, 3 * int(a.cap), tdt).view(3, -1)[:, dom].double()
        nn = device_view(a.n_neigh, int(a.cap), torch.int32)[dom]
        return F, nn

    def run(self, precision: int, dt: float, n_steps: int, use_graph: bool = True) -> None:
        L.call("ofp_ens_run", self.h, precision, dt, n_steps, int(use_graph), _stream())

    def test_masks(self, mask_cap: int = 0, pending: bool = True) -> int:
        """Test hook: replace the Verlet mask buffer by one of exactly mask_cap
        masks (0: keep it) and, with pending, start the next step with a rebuild.
        Returns the previous mask capacity."""
        old = C.c_size_t(0)
        L.call("ofp_ens_test_masks_", self.h, int(mask_cap), int(pending), C.byref(old))
        return int(old.value)

    def status(self) -> dict:
        st = L.EnsState()
        L.call("ofp_ens_status", self.h, C.byref(st), _stream())
        return {k: getattr(st, k) for k, _ in L.EnsState._fields_}

    # ------------------------------------------------------- inspection
    def gid(self) -> torch.Tensor:
        a = self.arrays()
        return _u32_tensor(a.gid, self.N)

    def lists(self):
        cl, vl = C.c_void_p(), C.c_void_p()
        L.call("ofp_ens_lists", self.h, C.byref(cl), C.byref(vl))
        cv = L.ClView()
        L.call("ofp_cl_get_view", cl, C.byref(cv))
        if not vl.value:  # cell-tile pair passes keep no Verlet list
            return cv, None
        vv = L.VlView()
        L.call("ofp_vl_get_view", vl, C.byref(vv))
        return cv, vv

    def cells(self):
        base = (C.c_uint32 * (self.S + 1))()
        divc = (C.c_uint32 * (3 * self.S))()
        lo = (C.c_double * (3 * self.S))()
        sp = (C.c_double * (3 * self.S))()
        L.call("ofp_ens_cells", self.h, base, divc, lo, sp)
        return (np.array(base[:], np.int64), np.array(divc[:], np.int64).reshape(-1, 3),
                np.array(lo[:]).reshape(-1, 3), np.array(sp[:]).reshape(-1, 3))

    def list_arrays(self) -> dict:
        """Host copies of the cell list and Verlet list of the last rebuild
        (global sorted indices; storage indices: domain [0, N), ghosts after)."""
        cv, vv = self.lists()
        n = self.N + int(self.status()["n_ghost"])
        out = {
            "cell_start": _u64_tensor(cv.cell_start, int(cv.n_cells) + 1).cpu().numpy(),
            "s2n": _u32_tensor(cv.sort_to_nonsort, n).cpu().numpy(),
            "dom": _u32_tensor(cv.dom_sort_ids, self.N).cpu().numpy(),
            "n_part": n,
        }
        if vv is not None:
            off = _u64_tensor(vv.offsets, self.N + 1).cpu().numpy()
            out["offsets"] = off
            out["nbr"] = _u32_tensor(vv.nbr, int(off[-1])).cpu().numpy()
        return out


def _u32_tensor(ptr: int, n: int) -> torch.Tensor:
    return device_view(ptr, n, torch.int32).clone()


def _u64_tensor(ptr: int, n: int) -> torch.Tensor:
    return device_view(ptr, n, torch.int64).clone()


def from_workloads(wls, dtype=torch.float32, opts: Options | None = None, mass: float = 1.0):
    """An ensemble of the `ensemble` workload's boxes (LJ sigma = 0.4 r_c)."""
    systems = [System(wl.n, wl.r_cut, wl.params["sigma"], wl.params.get("eps", 1.0), mass, tuple(wl.box_lo),
                      tuple(wl.box_hi)) for wl in wls]
    ens = Ensemble(systems, dtype, opts)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    pos = np.concatenate([wl.pos.astype(npdt) for wl in wls], axis=1)
    ens.set(torch.from_numpy(np.ascontiguousarray(pos)).cuda())
    return ens
