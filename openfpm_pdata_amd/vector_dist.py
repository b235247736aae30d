"""Host-side mirror of openfpm's ``vector_dist_gpu`` surface for the one-GPU
neighbour-search path, over libofp_nn.so.

This is the layer the tests and the benchmark drive; C++ users get the same
surface from ``include/openfpm_amd/vector_dist.hpp``.  Names and argument
meanings follow the reference (``src/Vector/vector_dist.hpp``):

* ``VectorDistGPU(n, box, bc, ghost)``          ctor ``vector_dist.hpp:498``
* ``map()``                                     ``:2158`` (one rank: periodic wrap)
* ``ghost_get(props, opt)``                     ``:2203``
* ``getCellListGPU(r_cut, opt, ...)``           ``:1222``
* ``updateCellListGPU(cl, props)``              ``:1334``
* ``restoreOrder(cl, props)``                   ``:3042``
* ``getVerletGPU(r_cut)``                       modelled on ``getVerlet`` ``:1514``

Device memory and streams come from PyTorch; every computation is a call into
the HIP library.  Positions are a ``(3, capacity)`` tensor (SoA planes, the
``memory_traits_inte`` layout of ``Point<3,St>``); each property is a
``(n_comp, capacity)`` tensor.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L

# ghost_get options: NO_POSITION, WITH_POSITION, NO_CHANGE_ELEMENTS
# (Vector/cuda/vector_dist_comm_util_funcs.cuh:13-15); SKIP_LABELLING and
# KEEP_PROPERTIES share one value (util/common_pdata.hpp:12-13): either name asks
# for the refresh of the cached ghost list
NO_POSITION = 1
WITH_POSITION = 2
NO_CHANGE_ELEMENTS = 4
SKIP_LABELLING = 512
KEEP_PROPERTIES = 512
RUN_ON_DEVICE = 1024

# cell-list options (vector_dist.hpp:1340-1372)
CL_NON_SYMMETRIC = L.OFP_CL_NON_SYMMETRIC
CL_GPU_REORDER_POSITION = L.OFP_CL_GPU_REORDER_POSITION
CL_GPU_REORDER_PROPERTY = L.OFP_CL_GPU_REORDER_PROPERTY
CL_GPU_REORDER = CL_GPU_REORDER_POSITION | CL_GPU_REORDER_PROPERTY
CL_QUANT = L.OFP_CL_QUANT

PERIODIC = 1
NON_PERIODIC = 0

# getCellListGPU's default ghostEnlargeFactor is a float (vector_dist.hpp:1223)
GHOST_ENLARGE_CL = float(np.float32(1.013))
# getVerlet magnifies the ghost by the double literal 1.013 (vector_dist.hpp:1537)
GHOST_ENLARGE_VL = 1.013


def _np_dtype(dt: torch.dtype):
    return {torch.float32: np.float32, torch.float64: np.float64}[dt]


def _ofp_dtype(dt: torch.dtype) -> int:
    return {torch.float32: L.OFP_F32, torch.float64: L.OFP_F64}[dt]


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _arr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def planes_of(t: torch.Tensor, rows=None) -> list:
    """ofp_plane descriptors of the rows of a (n_comp, capacity) tensor."""
    assert t.is_contiguous() and t.dim() == 2
    esz = t.element_size()
    rows = range(t.shape[0]) if rows is None else rows
    return [(t.data_ptr() + r * t.shape[1] * esz, esz) for r in rows]


def plane_array(pl: list):
    arr = (L.Plane * max(1, len(pl)))()
    for k, (ptr, esz) in enumerate(pl):
        arr[k].ptr = ptr
        arr[k].elem_bytes = esz
    return arr


class _DevArray:
    """Exposes a raw device pointer through __cuda_array_interface__."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {
            "shape": tuple(shape),
            "typestr": typestr,
            "data": (int(ptr), False),
            "version": 2,
            "strides": None,
        }


def device_view(ptr: int, n: int, dtype: torch.dtype) -> torch.Tensor:
    """Non-owning torch view of n elements at a library-owned device pointer."""
    typestr = {torch.int32: "<i4", torch.int64: "<i8", torch.float32: "<f4", torch.float64: "<f8"}[dtype]
    if n == 0:
        return torch.empty(0, dtype=dtype, device="cuda")
    return torch.as_tensor(_DevArray(ptr, (n,), typestr), device="cuda")


# ----------------------------------------------------------------- geometry
def shift_vectors(dtype: torch.dtype, lo, hi) -> np.ndarray:
    npd = _np_dtype(dtype)
    lo = np.asarray(lo, npd)
    hi = np.asarray(hi, npd)
    out = np.zeros((27, 3), npd)
    L.call("ofp_shift_vectors", _ofp_dtype(dtype), _arr(lo), _arr(hi), _arr(out))
    return out


def local_ghost_boxes(dtype: torch.dtype, lo, hi, g_lo, g_hi, bc):
    npd = _np_dtype(dtype)
    lo, hi, g_lo, g_hi = (np.asarray(v, npd) for v in (lo, hi, g_lo, g_hi))
    bc = np.asarray(bc, np.int32)
    blo = np.zeros((26, 3), npd)
    bhi = np.zeros((26, 3), npd)
    sv = np.zeros(26, np.uint32)
    nb = C.c_int(0)
    L.call("ofp_local_ghost_boxes", _ofp_dtype(dtype), _arr(lo), _arr(hi), _arr(g_lo), _arr(g_hi),
           _arr(bc), _arr(blo), _arr(bhi), _arr(sv), C.byref(nb))
    n = nb.value
    return blo[:n].copy(), bhi[:n].copy(), sv[:n].copy()


def cl_param(dtype: torch.dtype, plo, phi, r_cut: float, g_lo, g_hi, factor: float):
    npd = _np_dtype(dtype)
    plo, phi, g_lo, g_hi = (np.asarray(v, npd) for v in (plo, phi, g_lo, g_hi))
    blo = np.zeros(3, npd)
    bhi = np.zeros(3, npd)
    div = np.zeros(3, np.uint32)
    L.call("ofp_cl_param", _ofp_dtype(dtype), _arr(plo), _arr(phi), float(r_cut), _arr(g_lo), _arr(g_hi),
           float(factor), _arr(blo), _arr(bhi), _arr(div))
    return blo, bhi, div


# ---------------------------------------------------------------- cell list
class CellListGPU:
    """Owns one ofp_cl handle (CellList_gpu of the reference)."""

    def __init__(self, dtype: torch.dtype, box_lo, box_hi, div, opt: int, r_cut: float, box_nn: int = 1):
        npd = _np_dtype(dtype)
        self.dtype = dtype
        self.box_lo = np.asarray(box_lo, npd)
        self.box_hi = np.asarray(box_hi, npd)
        self.div = np.asarray(div, np.uint32)
        self.opt = opt
        self.r_cut = r_cut
        h = C.c_void_p()
        L.call("ofp_cl_create", C.byref(h), _ofp_dtype(dtype), _arr(self.box_lo), _arr(self.box_hi),
               _arr(self.div), opt)
        self.h = h
        self.box_nn = box_nn
        if box_nn != 1:
            L.call("ofp_cl_set_box_nn", h, box_nn)

    def __del__(self):
        if getattr(self, "h", None) and getattr(L, "_lib", None) is not None:
            L.lib().ofp_cl_destroy(self.h)
            self.h = None

    def construct(self, pos: torch.Tensor, n_tot: int, gm: int, pos_sorted=None, planes_in=(),
                  planes_out=()):
        pin = plane_array(list(planes_in))
        pout = plane_array(list(planes_out))
        L.call("ofp_cl_construct", self.h, pos.data_ptr(), pos.shape[1], n_tot, gm,
               pos_sorted.data_ptr() if pos_sorted is not None else None,
               pos_sorted.shape[1] if pos_sorted is not None else 0, pin, pout, len(planes_in), _stream())

    def view(self) -> L.ClView:
        v = L.ClView()
        L.call("ofp_cl_get_view", self.h, C.byref(v))
        return v

    # torch views of the library arrays (valid until the next construct)
    def arrays(self) -> dict:
        v = self.view()
        n = int(v.n_part)
        gm = int(v.ghost_marker)
        C_ = int(v.n_cells)
        cs = device_view(v.cell_start, C_ + 1, torch.int64)
        out = {
            "cell_start": (cs & 0xFFFFFFFF),
            "dom_start": (cs >> 32),
            "s2n": device_view(v.sort_to_nonsort, n, torch.int32),
            "n2s": device_view(v.nonsort_to_sort, n, torch.int32),
            "dom": device_view(v.dom_sort_ids, gm, torch.int32),
        }
        if v.quant:
            out["quant"] = device_view(v.quant, 4 * n, torch.int32).view(n, 4)
        return out

    def sparse(self) -> dict:
        """The sparse form of the last construct (CELLLIST_GPU_SPARSE): device
        views of the non-empty cells' ids, starts (n + 1) and stencil table
        (n, (2 box_nn + 1)^3; -1 empty)."""
        L.call("ofp_cl_build_sparse", self.h, _stream())
        v = L.ClSparseView()
        L.call("ofp_cl_get_sparse_view", self.h, C.byref(v))
        m, k = int(v.n_cells), int(v.nn_per_cell)
        return {"key": device_view(v.cell_key or 0, m, torch.int32),
                "start": device_view(v.cell_start, m + 1, torch.int32),
                "nn": device_view(v.cell_nn or 0, k * m, torch.int32).view(m, k),
                "box_nn": int(v.box_nn)}

    def restore(self, sorted_planes, orig_planes):
        L.call("ofp_cl_restore", self.h, plane_array(sorted_planes), plane_array(orig_planes),
               len(sorted_planes), _stream())

    def reorder(self, orig_planes, sorted_planes):
        L.call("ofp_cl_reorder", self.h, plane_array(orig_planes), plane_array(sorted_planes),
               len(orig_planes), _stream())


class VerletListGPU:
    def __init__(self):
        h = C.c_void_p()
        L.call("ofp_vl_create", C.byref(h))
        self.h = h
        self.cl = None

    def __del__(self):
        if getattr(self, "h", None) and getattr(L, "_lib", None) is not None:
            L.lib().ofp_vl_destroy(self.h)
            self.h = None

    def build(self, cl: CellListGPU, pos_sorted: torch.Tensor, r_cut: float):
        L.call("ofp_vl_build", self.h, cl.h, pos_sorted.data_ptr(), pos_sorted.shape[1], float(r_cut),
               _stream())
        self.cl = cl
        self.r_cut = r_cut

    def build_sym(self, full: "VerletListGPU", ghost_handle):
        """The symmetric half of `full` (getVerletSym): each interaction once."""
        L.call("ofp_vl_build_sym", self.h, full.h, ghost_handle, _stream())
        self.cl = full.cl
        self.r_cut = full.r_cut

    def view(self) -> L.VlView:
        v = L.VlView()
        L.call("ofp_vl_get_view", self.h, C.byref(v))
        return v

    def arrays(self):
        v = self.view()
        off = device_view(v.offsets, int(v.n_rows) + 1, torch.int64)
        nbr = device_view(v.nbr, int(v.n_entries), torch.int32)
        return off, nbr


@dataclass
class PropSpec:
    dtype: torch.dtype
    n_comp: int = 1


class VectorDistGPU:
    """vector_dist_gpu<3, St, aggregate<...>> on one GPU."""

    def __init__(self, n: int, box_lo, box_hi, bc, ghost: float, dtype=torch.float32,
                 props: dict | None = None, capacity: int | None = None, device="cuda"):
        self.dtype = dtype
        self.device = device
        npd = _np_dtype(dtype)
        self.box_lo = np.asarray(box_lo, npd)
        self.box_hi = np.asarray(box_hi, npd)
        self.bc = np.asarray(bc, np.int32)
        self.ghost_r = float(ghost)
        # Ghost<3,St> g(r): box [-r, r]
        self.g_lo = np.full(3, -npd(ghost), npd)
        self.g_hi = np.full(3, npd(ghost), npd)
        self.prop_spec = {k: (v if isinstance(v, PropSpec) else PropSpec(*v)) for k, v in (props or {}).items()}
        cap = max(16, capacity or int(n * 1.5) + 16)
        self._alloc(cap)
        self.size = n
        self.ghostMarker = n
        # decomposition geometry (one sub-domain == the domain)
        self.shifts = shift_vectors(dtype, self.box_lo, self.box_hi)
        self.blo, self.bhi, self.bsv = local_ghost_boxes(dtype, self.box_lo, self.box_hi, self.g_lo,
                                                         self.g_hi, self.bc)
        h = C.c_void_p()
        L.call("ofp_ghost_create", C.byref(h), _ofp_dtype(dtype), len(self.bsv), _arr(self.blo),
               _arr(self.bhi), _arr(self.bsv), _arr(self.shifts))
        self._gh = h
        self.sorted_by = None

    def __del__(self):
        if getattr(self, "_gh", None) and getattr(L, "_lib", None) is not None:
            L.lib().ofp_ghost_destroy(self._gh)
            self._gh = None

    # ------------------------------------------------------------ storage
    def _alloc(self, cap: int, keep: int = 0):
        old_pos = getattr(self, "pos", None)
        old_prp = getattr(self, "prp", {})
        self.cap = cap
        self.pos = torch.empty((3, cap), dtype=self.dtype, device=self.device)
        self.prp = {k: torch.empty((s.n_comp, cap), dtype=s.dtype, device=self.device)
                    for k, s in self.prop_spec.items()}
        if keep and old_pos is not None:
            self.pos[:, :keep] = old_pos[:, :keep]
            for k in self.prp:
                self.prp[k][:, :keep] = old_prp[k][:, :keep]
        self.pos_sorted = None
        self.prp_sorted = {}

    def reserve(self, n: int):
        if n > self.cap:
            self._alloc(max(n, int(self.cap * 1.5)), keep=self.size)

    def size_local(self) -> int:
        return self.ghostMarker

    def size_local_with_ghost(self) -> int:
        return self.size

    # ------------------------------------------------------ checkpoints
    def _io_props(self):
        return [(_np_dtype(s.dtype), s.n_comp) for s in self.prop_spec.values()]

    def save(self, filename: str) -> None:
        """save (vector_dist.hpp:2459-2463): the domain particles and all their
        properties, in the reference's HDF5 checkpoint layout (include/ofp_io.h)."""
        from . import h5io

        assert self.sorted_by is None, "restoreOrder before save"
        n = self.ghostMarker
        pos = self.pos[:, :n].cpu().numpy()
        arrs = [t[k, :n].cpu().numpy() for t in self.prp.values() for k in range(t.shape[0])]
        h5io.save(filename, pos, arrs, self._io_props())

    def load(self, filename: str) -> None:
        """load (vector_dist.hpp:2471-2475): the checkpoint's particles (every
        block, in file order) become the domain particles; no ghosts."""
        from . import h5io

        assert self.sorted_by is None, "restoreOrder before load"
        pos, arrs = h5io.load(filename, _np_dtype(self.dtype), self._io_props())
        n = pos.shape[1]
        self.size = 0
        self.reserve(n)
        self.pos[:, :n] = torch.from_numpy(np.ascontiguousarray(pos)).to(self.device)
        k = 0
        for t in self.prp.values():
            for c in range(t.shape[0]):
                t[c, :n] = torch.from_numpy(arrs[k]).to(self.device)
                k += 1
        self.size = n
        self.ghostMarker = n

    def getPosVector(self) -> torch.Tensor:
        return self.pos[:, : self.size]

    def getProp(self, name: str) -> torch.Tensor:
        return self.prp[name][:, : self.size]

    def _prop_planes(self, names, store=None):
        store = self.prp if store is None else store
        pl = []
        for nm in names:
            pl += planes_of(store[nm])
        return pl

    # -------------------------------------------------------------- map
    def map(self):
        """One-rank map(RUN_ON_DEVICE): drop ghosts, wrap periodic positions."""
        assert self.sorted_by is None, "restoreOrder before map"
        self.size = self.ghostMarker
        L.call("ofp_apply_bc", _ofp_dtype(self.dtype), self.pos.data_ptr(), self.cap, self.size,
               _arr(self.box_lo), _arr(self.box_hi), _arr(self.bc), _stream())

    # -------------------------------------------------------- ghost_get
    def ghost_get(self, props=(), opt: int = RUN_ON_DEVICE) -> int:
        """Local periodic ghosts appended after ghostMarker; returns G."""
        assert self.sorted_by is None, "restoreOrder before ghost_get"
        n_dom = self.ghostMarker
        planes = self._prop_planes(props)
        if opt & SKIP_LABELLING:
            with_pos = 0 if (opt & NO_POSITION) else 1
            L.call("ofp_ghost_refresh", self._gh, self.pos.data_ptr(), self.cap, n_dom, with_pos,
                   plane_array(planes), len(planes), _stream())
            return self.size - n_dom
        self.size = n_dom
        g = C.c_size_t(0)
        L.call("ofp_ghost_label", self._gh, self.pos.data_ptr(), self.cap, n_dom, C.byref(g), _stream())
        G = g.value
        if n_dom + G > self.cap:
            self.reserve(n_dom + G)
            planes = self._prop_planes(props)
        L.call("ofp_ghost_emit", self._gh, self.pos.data_ptr(), self.cap, n_dom, n_dom, plane_array(planes),
               len(planes), _stream())
        self.size = n_dom + G
        return G

    def ghost_opart(self):
        pid = C.c_void_p()
        sid = C.c_void_p()
        n = C.c_size_t()
        L.call("ofp_ghost_opart", self._gh, C.byref(pid), C.byref(sid), C.byref(n))
        return (device_view(pid.value or 0, n.value, torch.int32),
                device_view(sid.value or 0, n.value, torch.int32))

    # -------------------------------------------------------- cell lists
    def getCellListGPU(self, r_cut: float, opt: int = CL_NON_SYMMETRIC, NNIteratorBox: int = 1,
                       ghostEnlargeFactor: float = GHOST_ENLARGE_CL) -> CellListGPU:
        """Cells of size >= r_cut; the neighbour stencil spans NNIteratorBox cells
        on each side (setBoxNN, vector_dist.hpp:1252)."""
        blo, bhi, div = cl_param(self.dtype, self.box_lo, self.box_hi, r_cut, self.g_lo, self.g_hi,
                                 ghostEnlargeFactor)
        return CellListGPU(self.dtype, blo, bhi, div, opt, r_cut, NNIteratorBox)

    def updateCellListGPU(self, cl: CellListGPU, props=()):
        """Build; with CL_GPU_REORDER_* the vector is left in the sorted order."""
        reorder_pos = bool(cl.opt & CL_GPU_REORDER_POSITION)
        reorder_prp = bool(cl.opt & CL_GPU_REORDER_PROPERTY)
        if reorder_pos and (self.pos_sorted is None or self.pos_sorted.shape[1] != self.cap):
            self.pos_sorted = torch.empty_like(self.pos)
        pin, pout = [], []
        if reorder_prp:
            for nm in props:
                if nm not in self.prp_sorted or self.prp_sorted[nm].shape[1] != self.cap:
                    self.prp_sorted[nm] = torch.empty_like(self.prp[nm])
            pin = self._prop_planes(props)
            pout = self._prop_planes(props, self.prp_sorted)
        cl.construct(self.pos, self.size, self.ghostMarker, self.pos_sorted if reorder_pos else None, pin, pout)
        if reorder_pos:
            self.pos, self.pos_sorted = self.pos_sorted, self.pos
        if reorder_prp:
            for nm in props:
                self.prp[nm], self.prp_sorted[nm] = self.prp_sorted[nm], self.prp[nm]
        if reorder_pos or reorder_prp:
            self.sorted_by = cl
        return cl

    def restoreOrder(self, cl: CellListGPU, props=()):
        """Scatter positions (if reordered) and props back to the original order."""
        assert self.sorted_by is cl, "vector is not in this cell list's order"
        if cl.opt & CL_GPU_REORDER_POSITION:
            cl.restore(planes_of(self.pos), planes_of(self.pos_sorted))
            self.pos, self.pos_sorted = self.pos_sorted, self.pos
        for nm in props:
            if nm not in self.prp_sorted or self.prp_sorted[nm].shape[1] != self.cap:
                self.prp_sorted[nm] = torch.empty_like(self.prp[nm])
            cl.restore(planes_of(self.prp[nm]), planes_of(self.prp_sorted[nm]))
            self.prp[nm], self.prp_sorted[nm] = self.prp_sorted[nm], self.prp[nm]
        self.sorted_by = None

    def ghost_sync_sorted(self, cl: CellListGPU, props):
        """ghost_get of props while the vector is in cl's sorted order."""
        planes = self._prop_planes(props)
        L.call("ofp_ghost_sync_sorted", self._gh, cl.h, plane_array(planes), len(planes), _stream())

    def ghost_put_sorted(self, cl: CellListGPU, planes):
        """ghost_put<add_> while the vector is in cl's sorted order: planes is a
        list of device tensors of n_part values (float32, float64, int32 or int64,
        sorted index); each ghost's value is added to its source particle's."""
        types = {torch.float32: L.OFP_PUT_F32, torch.float64: L.OFP_PUT_F64, torch.int32: L.OFP_PUT_U32,
                 torch.int64: L.OFP_PUT_I64}
        check_sorted_planes(planes, int(cl.view().n_part), tuple(types), "ghost_put_sorted")
        ptrs = (C.c_void_p * max(1, len(planes)))(*[t.data_ptr() for t in planes])
        tys = (C.c_int * max(1, len(planes)))(*[types[t.dtype] for t in planes])
        L.call("ofp_ghost_put_sorted", self._gh, cl.h, ptrs, tys, len(planes), _stream())

    # ------------------------------------------------------------ Verlet
    def getVerletSymGPU(self, r_cut: float, opt: int = CL_NON_SYMMETRIC, sub: int = 1):
        """getVerletSym (vector_dist.hpp:1412-1416): the full list of getVerletGPU
        and its symmetric half; returns (cl, full, sym)."""
        cl, full = self.getVerletGPU(r_cut, opt, sub)
        sym = VerletListGPU()
        sym.build_sym(full, self._gh)
        return cl, full, sym

    def getVerletGPU(self, r_cut: float, opt: int = CL_NON_SYMMETRIC, sub: int = 1):
        """Verlet list of the current particles: rows in domain-sort order of an
        internal cell list over the box enlarged by the ghost x 1.013 (getVerlet,
        vector_dist.hpp:1514-1545), with cells of size >= r_cut / sub and a
        stencil of `sub` cells each side.  The vector is reordered into that cell
        list's order (call restoreOrder to undo)."""
        cl = self.getCellListGPU(r_cut / sub, opt | CL_GPU_REORDER_POSITION, NNIteratorBox=sub,
                                 ghostEnlargeFactor=GHOST_ENLARGE_VL)
        self.updateCellListGPU(cl)
        vl = VerletListGPU()
        vl.build(cl, self.pos, r_cut)
        return cl, vl


def check_sorted_planes(planes, n: int, dtypes, what: str) -> None:
    """Planes a kernel indexes by sorted index up to n: each a contiguous device
    tensor of one of dtypes with at least n values (a short, strided or host
    tensor is rejected here instead of being written out of bounds)."""
    for k, t in enumerate(planes):
        if not isinstance(t, torch.Tensor) or t.dtype not in dtypes:
            raise ValueError(f"{what}: plane {k} must be a tensor of one of {dtypes}")
        if not t.is_contiguous():
            raise ValueError(f"{what}: plane {k} is not contiguous")
        if t.numel() < n:
            raise ValueError(f"{what}: plane {k} holds {t.numel()} values, fewer than n_part = {n}")
        if not t.is_cuda:
            raise ValueError(f"{what}: plane {k} is not on the GPU")


def check_pos_sorted(pos: torch.Tensor, n: int, dtype, what: str) -> None:
    """Sorted positions read up to n_part: a (3, >= n) tensor of the list's dtype
    with contiguous planes, on the GPU."""
    if not isinstance(pos, torch.Tensor) or pos.dim() != 2 or pos.shape[0] != 3 or pos.dtype != dtype:
        raise ValueError(f"{what}: positions must be a (3, n) tensor of {dtype}")
    if pos.stride(1) != 1:
        raise ValueError(f"{what}: position planes are not contiguous")
    if pos.shape[1] < n:
        raise ValueError(f"{what}: positions hold {pos.shape[1]} particles, fewer than n_part = {n}")
    if not pos.is_cuda:
        raise ValueError(f"{what}: positions are not on the GPU")


def pair_lj_sym(cl: CellListGPU, vl: VerletListGPU, pos_sorted: torch.Tensor, r_cut: float, sigma: float,
                eps: float, counts: bool = True):
    """LJ over a symmetric list: (forces (3, n_part) float64, neighbour counts
    (n_part,) int32 or None), both ends of each pair, ghosts included, sorted
    index."""
    n = int(cl.view().n_part)
    want = torch.float64 if cl.view().dtype == L.OFP_F64 else torch.float32
    check_pos_sorted(pos_sorted, n, want, "pair_lj_sym")
    F = torch.empty((3, max(1, n)), dtype=torch.float64, device=pos_sorted.device)
    nn = torch.empty(max(1, n), dtype=torch.int32, device=pos_sorted.device) if counts else None
    prm = L.LJParams(r_cut, sigma, eps)
    L.call("ofp_pair_lj_sym", cl.h, vl.h, pos_sorted.data_ptr(), pos_sorted.shape[1], C.byref(prm),
           F[0].data_ptr(), F[1].data_ptr(), F[2].data_ptr(), nn.data_ptr() if counts else None, _stream())
    return F, nn
