"""Remote ghosts: the send side of ghost_get on one rank of a multi-rank run,
over the C ABI (ofp_ghost_remote_*).

GhostRemote mirrors the reference's per-rank steps: labelParticlesGhost_impl
(Vector/cuda/vector_dist_comm_util_funcs.cuh:53-177: label every particle
against the internal ghost boxes other ranks need, order the outputs by
processor, count them per processor) and the send-buffer fill
process_ghost_particles_pos / _prp (Vector/cuda/vector_dist_cuda_funcs.cuh:90-131).

regular_decomposition builds the boxes for one rank of a regular px x py x pz
decomposition of the box, the role CartDecomposition plays for the reference
(Decomposition/CartDecomposition.hpp, ie_ghost's internal ghost boxes).

Map mirrors map(RUN_ON_DEVICE) on one rank (Vector/vector_dist_comm.hpp:
1398-1456 and fill_send_map_buf, :1240-1310): wrap, find every particle's
processor among the sub-domains (regular_subdomains builds them for a regular
decomposition), order by destination, and split into the particles that stay
and the send buffer.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .vector_dist import _arr, _np_dtype, _ofp_dtype, _stream, device_view, plane_array, shift_vectors


def regular_decomposition(dtype: torch.dtype, box_lo, box_hi, bc, dims, rank: int, ghost: float):
    """Sub-domain of `rank` and the internal ghost boxes it sends to the others.

    The box [lo, hi) is cut into dims[k] equal slabs per dimension; rank
    r = i + px (j + py k) owns D_r = prod_k [lo_k + c_k d_k, lo_k + (c_k + 1) d_k)
    (the last slab ends at hi_k).  For every other rank q and every shift s of
    the 27 (ofp_shift_vectors order; a ghost sits at x - s), the box
    D_r n (E_q + s), with E_q = D_q enlarged by `ghost` on every side, holds the
    particles of r that q needs as ghosts at x - s.  Boxes of positive volume
    only, in (q, shift id) order; in a non-periodic dimension only shifts with a
    zero component there.  Corners are computed in double and stored in the
    positions' dtype.

    Returns (dom_lo, dom_hi, box_lo (nb, 3), box_hi (nb, 3), proc (nb,),
    shift_id (nb,), shifts (27, 3)).
    """
    npd = _np_dtype(dtype)
    lo = np.asarray(box_lo, np.float64)
    hi = np.asarray(box_hi, np.float64)
    dims = [int(d) for d in dims]
    assert all(d >= 1 for d in dims) and 0 <= rank < dims[0] * dims[1] * dims[2]
    d = (hi - lo) / np.asarray(dims, np.float64)

    def coords(r):
        return (r % dims[0], (r // dims[0]) % dims[1], r // (dims[0] * dims[1]))

    def dom(r):
        c = coords(r)
        a = np.array([lo[k] + c[k] * d[k] for k in range(3)])
        b = np.array([hi[k] if c[k] == dims[k] - 1 else lo[k] + (c[k] + 1) * d[k] for k in range(3)])
        return a, b

    shifts = shift_vectors(dtype, box_lo, box_hi)
    sh64 = shifts.astype(np.float64)
    r_lo, r_hi = dom(rank)
    blo, bhi, proc, sid = [], [], [], []
    for q in range(dims[0] * dims[1] * dims[2]):
        if q == rank:
            continue
        q_lo, q_hi = dom(q)
        e_lo, e_hi = q_lo - ghost, q_hi + ghost
        for s in range(27):
            if any(sh64[s, k] != 0.0 and not bc[k] for k in range(3)):
                continue
            a = np.maximum(r_lo, e_lo + sh64[s])
            b = np.minimum(r_hi, e_hi + sh64[s])
            if np.all(b > a):
                blo.append(a)
                bhi.append(b)
                proc.append(q)
                sid.append(s)
    nb = len(blo)
    return (r_lo.astype(npd), r_hi.astype(npd), np.asarray(blo, npd).reshape(nb, 3),
            np.asarray(bhi, npd).reshape(nb, 3), np.asarray(proc, np.uint32), np.asarray(sid, np.uint32), shifts)


def regular_subdomains(dtype: torch.dtype, box_lo, box_hi, dims):
    """The sub-domains of a regular px x py x pz decomposition of [lo, hi) (the
    dom() of regular_decomposition for every rank r = i + px (j + py k)):
    (sub_lo (nr, 3), sub_hi (nr, 3), proc (nr,)), corners in the dtype."""
    npd = _np_dtype(dtype)
    lo = np.asarray(box_lo, np.float64)
    hi = np.asarray(box_hi, np.float64)
    dims = [int(d) for d in dims]
    d = (hi - lo) / np.asarray(dims, np.float64)
    nr = dims[0] * dims[1] * dims[2]
    slo = np.zeros((nr, 3))
    shi = np.zeros((nr, 3))
    for r in range(nr):
        c = (r % dims[0], (r // dims[0]) % dims[1], r // (dims[0] * dims[1]))
        for k in range(3):
            slo[r, k] = lo[k] + c[k] * d[k]
            shi[r, k] = hi[k] if c[k] == dims[k] - 1 else lo[k] + (c[k] + 1) * d[k]
    return slo.astype(npd), shi.astype(npd), np.arange(nr, dtype=np.uint32)


def geo(box_lo, box_hi):
    """The geometric cell list the library builds for these boxes (host only):
    (grid_lo, cell_size, div, cell_start, cell_boxes)."""
    blo = np.ascontiguousarray(box_lo, np.float64).reshape(-1, 3)
    bhi = np.ascontiguousarray(box_hi, np.float64).reshape(-1, 3)
    nb = blo.shape[0]
    glo = np.zeros(3)
    cs = np.zeros(3)
    div = np.zeros(3, np.uint32)
    ne = C.c_size_t(0)
    L.call("ofp_ghost_remote_geo", nb, _arr(blo), _arr(bhi), _arr(glo), _arr(cs), _arr(div), None, None,
           C.byref(ne))
    start = np.zeros(int(np.prod(div.astype(np.int64))) + 1, np.uint32)
    boxes = np.zeros(max(1, ne.value), np.uint32)
    L.call("ofp_ghost_remote_geo", nb, _arr(blo), _arr(bhi), _arr(glo), _arr(cs), _arr(div), _arr(start),
           _arr(boxes), C.byref(ne))
    return glo, cs, div, start, boxes[: ne.value]


class GhostRemote:
    """One rank's internal ghost boxes for the other ranks, and the last label."""

    def __init__(self, dtype: torch.dtype, box_lo, box_hi, proc, shift_id, shifts):
        npd = _np_dtype(dtype)
        self.dtype = dtype
        self._blo = np.ascontiguousarray(box_lo, npd).reshape(-1, 3)
        self._bhi = np.ascontiguousarray(box_hi, npd).reshape(-1, 3)
        self._proc = np.ascontiguousarray(proc, np.uint32)
        self._sid = np.ascontiguousarray(shift_id, np.uint32)
        self._shifts = np.ascontiguousarray(shifts, npd).reshape(-1, 3)
        nb = self._blo.shape[0]
        h = C.c_void_p()
        L.call("ofp_ghost_remote_create", C.byref(h), _ofp_dtype(dtype), nb, _arr(self._blo), _arr(self._bhi),
               _arr(self._proc), _arr(self._sid), self._shifts.shape[0], _arr(self._shifts))
        self.h = h
        self.n_out = 0
        self.n_procs = 0

    def label(self, pos: torch.Tensor, n: int):
        """Labels particles [0, n) of pos (3, capacity); returns (outputs, processors)."""
        assert pos.dtype == self.dtype and pos.is_cuda and pos.is_contiguous()
        n_out = C.c_size_t(0)
        n_procs = C.c_int(0)
        L.call("ofp_ghost_remote_label", self.h, pos.data_ptr(), pos.shape[1], n, C.byref(n_out),
               C.byref(n_procs), _stream())
        self.n_out, self.n_procs = n_out.value, n_procs.value
        return self.n_out, self.n_procs

    def procs(self):
        """(destination processors ascending, output count of each)."""
        p = np.zeros(max(1, self.n_procs), np.uint32)
        c = np.zeros(max(1, self.n_procs), np.uint64)
        L.call("ofp_ghost_remote_procs", self.h, _arr(p), _arr(c))
        return p[: self.n_procs], c[: self.n_procs]

    def opart(self):
        """Device views of the ordered outputs: processor (int32) and particle | shift << 32 (int64)."""
        pp = C.c_void_p()
        ip = C.c_void_p()
        n = C.c_size_t(0)
        L.call("ofp_ghost_remote_opart", self.h, C.byref(pp), C.byref(ip), C.byref(n))
        return device_view(pp.value or 0, n.value, torch.int32), device_view(ip.value or 0, n.value, torch.int64)

    def pack(self, pos: torch.Tensor, send_pos: torch.Tensor, planes_in=(), planes_out=()):
        """Send buffer: send_pos[:, i] = pos[:, id_i] - shift[sid_i], planes copied."""
        assert send_pos.dtype == self.dtype and send_pos.shape[1] >= self.n_out
        pin = plane_array(list(planes_in))
        pout = plane_array(list(planes_out))
        L.call("ofp_ghost_remote_pack", self.h, pos.data_ptr(), pos.shape[1], send_pos.data_ptr(),
               send_pos.shape[1], pin, pout, len(planes_in), _stream())

    def close(self):
        if getattr(self, "h", None):
            L.call("ofp_ghost_remote_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Map:
    """One rank's map: the sub-domains of every rank, the domain, and the last
    label (staying count, destinations, lost count)."""

    def __init__(self, dtype: torch.dtype, sub_lo, sub_hi, sub_proc, dom_lo, dom_hi, bc, rank: int):
        npd = _np_dtype(dtype)
        self.dtype = dtype
        self.rank = int(rank)
        self._slo = np.ascontiguousarray(sub_lo, npd).reshape(-1, 3)
        self._shi = np.ascontiguousarray(sub_hi, npd).reshape(-1, 3)
        self._sp = np.ascontiguousarray(sub_proc, np.uint32)
        self._dlo = np.ascontiguousarray(dom_lo, npd)
        self._dhi = np.ascontiguousarray(dom_hi, npd)
        self._bc = np.ascontiguousarray(bc, np.int32)
        h = C.c_void_p()
        L.call("ofp_map_create", C.byref(h), _ofp_dtype(dtype), self._slo.shape[0], _arr(self._slo), _arr(self._shi),
               _arr(self._sp), _arr(self._dlo), _arr(self._dhi), _arr(self._bc), self.rank)
        self.h = h
        self.n = self.n_stay = self.n_dest = self.n_lost = 0

    def label(self, pos: torch.Tensor, n: int):
        """Wraps particles [0, n) of pos (3, capacity) in place and labels them;
        returns (staying, destinations, lost)."""
        assert pos.dtype == self.dtype and pos.is_cuda and pos.is_contiguous()
        ns, nd, nl = C.c_size_t(0), C.c_int(0), C.c_size_t(0)
        L.call("ofp_map_label", self.h, pos.data_ptr(), pos.shape[1], n, C.byref(ns), C.byref(nd), C.byref(nl),
               _stream())
        self.n, self.n_stay, self.n_dest, self.n_lost = n, ns.value, nd.value, nl.value
        return self.n_stay, self.n_dest, self.n_lost

    @property
    def n_send(self) -> int:
        return self.n - self.n_stay - self.n_lost

    def procs(self):
        """(destination processors other than this rank, ascending; particle count of each)."""
        p = np.zeros(max(1, self.n_dest), np.uint32)
        c = np.zeros(max(1, self.n_dest), np.uint64)
        L.call("ofp_map_procs", self.h, _arr(p), _arr(c))
        return p[: self.n_dest], c[: self.n_dest]

    def pack(self, pos: torch.Tensor, stay_pos: torch.Tensor, send_pos: torch.Tensor, planes_in=(), stay_planes=(),
             send_planes=()):
        """stay_pos / send_pos (3, >= count) and their planes from pos and planes_in."""
        assert stay_pos.dtype == self.dtype and send_pos.dtype == self.dtype
        assert stay_pos.shape[1] >= self.n_stay and send_pos.shape[1] >= self.n_send
        pin = plane_array(list(planes_in))
        pst = plane_array(list(stay_planes))
        psd = plane_array(list(send_planes))
        L.call("ofp_map_pack", self.h, pos.data_ptr(), pos.shape[1], stay_pos.data_ptr(), stay_pos.shape[1],
               send_pos.data_ptr(), send_pos.shape[1], pin, pst, psd, len(planes_in), _stream())

    def close(self):
        if getattr(self, "h", None):
            L.call("ofp_map_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
