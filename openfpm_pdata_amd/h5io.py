"""vector_dist save / load in the reference's HDF5 checkpoint layout (ctypes
over libofp_io.so; the layout is described in include/ofp_io.h).

Reference: vector_dist::save / load (src/Vector/vector_dist.hpp:2459-2475).
Host arrays only; VectorDistGPU.save / load copy between device and host.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libofp_io.so")
_io = None


class Field(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("elem_bytes", C.c_uint32), ("offset", C.c_uint32)]


def lib():
    global _io
    if _io is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C openfpm_pdata_amd/csrc`")
        io = C.CDLL(LIB_PATH)
        io.ofp_io_last_error.restype = C.c_char_p
        io.ofp_h5_save.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_size_t, C.c_size_t, C.POINTER(Field),
                                   C.c_int, C.c_uint32]
        io.ofp_h5_load_count.argtypes = [C.c_char_p, C.c_int, C.c_uint32, C.POINTER(C.c_size_t)]
        io.ofp_h5_load.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_size_t, C.c_size_t, C.POINTER(Field),
                                   C.c_int, C.c_uint32, C.POINTER(C.c_size_t)]
        _io = io
    return _io


def _check(rc: int) -> None:
    if rc != 0:
        raise L.OfpError(rc, lib().ofp_io_last_error().decode())


def layout(props):
    """Packed property object of an aggregate: props = [(numpy dtype, n
    components), ...] in declaration order; each member at its natural alignment,
    the object padded to the largest one.  Returns ([(offset, elem_bytes) per
    component plane], object bytes)."""
    off, planes, align = 0, [], 1
    for dt, nc in props:
        e = np.dtype(dt).itemsize
        off = (off + e - 1) // e * e
        for k in range(nc):
            planes.append((off + k * e, e))
        off += nc * e
        align = max(align, e)
    return planes, (off + align - 1) // align * align


def _fields(arrays, planes):
    assert len(arrays) == len(planes)
    fs = (Field * max(1, len(arrays)))()
    for k, (a, (o, e)) in enumerate(zip(arrays, planes)):
        assert a.flags["C_CONTIGUOUS"] and a.dtype.itemsize == e
        fs[k] = Field(a.ctypes.data, e, o)
    return fs


def save(filename: str, pos: np.ndarray, prop_arrays, props) -> None:
    """pos: (3, n) float32/float64; prop_arrays: one 1-D array per component
    plane (in the order of layout(props))."""
    pos = np.ascontiguousarray(pos)
    dt = L.OFP_F64 if pos.dtype == np.float64 else L.OFP_F32
    planes, obj = layout(props)
    arrs = [np.ascontiguousarray(a) for a in prop_arrays]
    fs = _fields(arrs, planes)
    _check(lib().ofp_h5_save(os.fsencode(filename), dt, pos.ctypes.data, pos.shape[1], pos.shape[1], fs,
                             len(arrs), obj))


def load(filename: str, dtype, props):
    """(pos (3, n), [component planes]) of a checkpoint whose layout is (dtype,
    props)."""
    dtype = np.dtype(dtype)
    dt = L.OFP_F64 if dtype == np.float64 else L.OFP_F32
    planes, obj = layout(props)
    n = C.c_size_t(0)
    _check(lib().ofp_h5_load_count(os.fsencode(filename), dt, obj, C.byref(n)))
    m = n.value
    pos = np.zeros((3, max(1, m)), dtype)
    arrs = []
    for dtp, nc in props:
        arrs += [np.zeros(max(1, m), dtp) for _ in range(nc)]
    fs = _fields(arrs, planes)
    _check(lib().ofp_h5_load(os.fsencode(filename), dt, pos.ctypes.data, pos.shape[1], m, fs, len(arrs), obj,
                             C.byref(n)))
    return pos[:, :m], [a[:m] for a in arrs]
