"""ctypes binding of libofp_nn.so (the C ABI declared in include/ofp_nn.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C openfpm_pdata_amd/csrc``).  Loading fails loudly when it is missing:
there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libofp_nn.so")

OFP_OK = 0
OFP_E_ARG = -1
OFP_E_HIP = -2
OFP_E_CAPACITY = -3
OFP_E_LIMIT = -4
OFP_E_STATE = -5

OFP_F32 = 0
OFP_F64 = 1
OFP_PREC_F64 = 0
OFP_PREC_F32C = 1
OFP_PUT_F32 = 0
OFP_PUT_F64 = 1
OFP_PUT_U32 = 2
OFP_PUT_I64 = 3

OFP_CL_NON_SYMMETRIC = 0x1
OFP_CL_GPU_REORDER_POSITION = 0x100
OFP_CL_GPU_REORDER_PROPERTY = 0x200
OFP_CL_QUANT = 0x10000

OFP_RED_SUM = 0
OFP_RED_MAX = 1
OFP_RED_MIN = 2


class OfpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ofp error {code}: {msg}")
        self.code = code


class Plane(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("elem_bytes", C.c_uint32), ("pad_", C.c_uint32)]


class ClView(C.Structure):
    _fields_ = [
        ("cell_start", C.c_void_p),
        ("sort_to_nonsort", C.c_void_p),
        ("nonsort_to_sort", C.c_void_p),
        ("dom_sort_ids", C.c_void_p),
        ("quant", C.c_void_p),
        ("lo", C.c_double * 3),
        ("spacing", C.c_double * 3),
        ("div", C.c_uint32 * 3),
        ("divc", C.c_uint32 * 3),
        ("n_cells", C.c_uint64),
        ("n_part", C.c_uint64),
        ("ghost_marker", C.c_uint64),
        ("dtype", C.c_int),
        ("pad_", C.c_int),
    ]


class VlView(C.Structure):
    _fields_ = [
        ("offsets", C.c_void_p),
        ("nbr", C.c_void_p),
        ("n_rows", C.c_uint64),
        ("n_entries", C.c_uint64),
    ]


class ClSparseView(C.Structure):
    _fields_ = [("cell_key", C.c_void_p), ("cell_start", C.c_void_p), ("cell_nn", C.c_void_p),
                ("n_cells", C.c_uint64), ("box_nn", C.c_uint32), ("nn_per_cell", C.c_uint32)]


class LJParams(C.Structure):
    _fields_ = [("r_cut", C.c_double), ("sigma", C.c_double), ("eps", C.c_double)]


class SPHParams(C.Structure):
    _fields_ = [("h", C.c_double), ("mass", C.c_double), ("rho0", C.c_double), ("c0", C.c_double)]


class SysDesc(C.Structure):
    _fields_ = [("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3), ("r_cut", C.c_double),
                ("sigma", C.c_double), ("eps", C.c_double), ("mass", C.c_double), ("n", C.c_uint64)]


class EnsOpts(C.Structure):
    _fields_ = [("skin", C.c_double), ("sub", C.c_uint32), ("lanes", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("verlet_rows", C.c_uint32), ("headroom", C.c_double), ("pair_walk", C.c_uint32),
                ("split_kick", C.c_uint32)]


class EnsState(C.Structure):
    _fields_ = [("n_dom", C.c_uint64), ("n_ghost", C.c_uint64), ("n_entries", C.c_uint64),
                ("ghost_cap", C.c_uint64), ("entry_cap", C.c_uint64), ("rebuilds", C.c_uint64),
                ("steps", C.c_uint64), ("recaptures", C.c_uint64), ("rebuild_pending", C.c_uint32),
                ("overflow", C.c_uint32)]


class EnsArrays(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("vel", C.c_void_p), ("force", C.c_void_p), ("gid", C.c_void_p),
                ("sid", C.c_void_p), ("force_sorted", C.c_void_p), ("n_neigh", C.c_void_p), ("cap", C.c_uint64),
                ("n_dom", C.c_uint64)]


_vp = C.c_void_p
_sz = C.c_size_t
_i = C.c_int
_u32p = C.POINTER(C.c_uint32)

# name -> (restype, argtypes)
SIGNATURES = {
    "ofp_last_error": (C.c_char_p, []),
    "ofp_abi_version": (_i, []),
    "ofp_device_info": (_i, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, _sz]),
    "ofp_shift_vectors": (_i, [_i, _vp, _vp, _vp]),
    "ofp_local_ghost_boxes": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_int)]),
    "ofp_cl_param": (_i, [_i, _vp, _vp, C.c_double, _vp, _vp, C.c_double, _vp, _vp, _vp]),
    "ofp_ghost_create": (_i, [C.POINTER(_vp), _i, _i, _vp, _vp, _vp, _vp]),
    "ofp_ghost_destroy": (_i, [_vp]),
    "ofp_ghost_label": (_i, [_vp, _vp, _sz, _sz, C.POINTER(_sz), _vp]),
    "ofp_ghost_emit": (_i, [_vp, _vp, _sz, _sz, _sz, _vp, _i, _vp]),
    "ofp_ghost_refresh": (_i, [_vp, _vp, _sz, _sz, _i, _vp, _i, _vp]),
    "ofp_ghost_opart": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_sz)]),
    "ofp_ghost_sync_sorted": (_i, [_vp, _vp, _vp, _i, _vp]),
    "ofp_ghost_remote_create": (_i, [C.POINTER(_vp), _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "ofp_ghost_remote_destroy": (_i, [_vp]),
    "ofp_ghost_remote_geo": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_sz)]),
    "ofp_ghost_remote_label": (_i, [_vp, _vp, _sz, _sz, C.POINTER(_sz), C.POINTER(_i), _vp]),
    "ofp_ghost_remote_procs": (_i, [_vp, _vp, _vp]),
    "ofp_ghost_remote_opart": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_sz)]),
    "ofp_ghost_remote_pack": (_i, [_vp, _vp, _sz, _vp, _sz, _vp, _vp, _i, _vp]),
    "ofp_cl_build_sparse": (_i, [_vp, _vp]),
    "ofp_cl_get_sparse_view": (_i, [_vp, C.POINTER(ClSparseView)]),
    "ofp_vl_build_sym": (_i, [_vp, _vp, _vp, _vp]),
    "ofp_pair_lj_sym": (_i, [_vp, _vp, _vp, _sz, C.POINTER(LJParams), _vp, _vp, _vp, _vp, _vp]),
    "ofp_ghost_put_sorted": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "ofp_ghost_put": (_i, [_vp, _vp, _vp, _i, _vp]),
    "ofp_map_create": (_i, [C.POINTER(_vp), _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32]),
    "ofp_map_destroy": (_i, [_vp]),
    "ofp_map_label": (_i, [_vp, _vp, _sz, _sz, C.POINTER(_sz), C.POINTER(_i), C.POINTER(_sz), _vp]),
    "ofp_map_procs": (_i, [_vp, _vp, _vp]),
    "ofp_map_pack": (_i, [_vp, _vp, _sz, _vp, _sz, _vp, _sz, _vp, _vp, _vp, _i, _vp]),
    "ofp_apply_bc": (_i, [_i, _vp, _sz, _sz, _vp, _vp, _vp, _vp]),
    "ofp_cl_create": (_i, [C.POINTER(_vp), _i, _vp, _vp, _vp, C.c_uint32]),
    "ofp_cl_destroy": (_i, [_vp]),
    "ofp_cl_construct": (_i, [_vp, _vp, _sz, _sz, _sz, _vp, _sz, _vp, _vp, _i, _vp]),
    "ofp_cl_get_view": (_i, [_vp, C.POINTER(ClView)]),
    "ofp_cl_set_box_nn": (_i, [_vp, C.c_uint32]),
    "ofp_cl_restore": (_i, [_vp, _vp, _vp, _i, _vp]),
    "ofp_cl_reorder": (_i, [_vp, _vp, _vp, _i, _vp]),
    "ofp_vl_create": (_i, [C.POINTER(_vp)]),
    "ofp_vl_destroy": (_i, [_vp]),
    "ofp_vl_build": (_i, [_vp, _vp, _vp, _sz, C.c_double, _vp]),
    "ofp_vl_get_view": (_i, [_vp, C.POINTER(VlView)]),
    "ofp_pair_lj": (_i, [_vp, _vp, _vp, _sz, _i, C.POINTER(LJParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ofp_sph_density": (_i, [_vp, _vp, _vp, _sz, _i, C.POINTER(SPHParams), _vp, _vp]),
    "ofp_sph_eos": (_i, [_i, _vp, _sz, C.POINTER(SPHParams), _vp, _vp, _vp]),
    "ofp_sph_force": (_i, [_vp, _vp, _vp, _sz, _i, C.POINTER(SPHParams), _vp, _vp, _vp, _vp, _vp]),
    "ofp_reduce": (_i, [_i, _vp, _vp, _sz, _i, C.POINTER(C.c_double), _vp]),
    "ofp_reduce_noalloc": (_i, [_i, _vp, _vp, _sz, _i, _vp, _vp]),
    "ofp_ens_create": (_i, [C.POINTER(_vp), _i, C.c_uint32, C.POINTER(SysDesc), C.POINTER(EnsOpts)]),
    "ofp_ens_destroy": (_i, [_vp]),
    "ofp_ens_first": (_i, [_vp, C.POINTER(C.c_uint64)]),
    "ofp_ens_arrays_get": (_i, [_vp, C.POINTER(EnsArrays)]),
    "ofp_ens_lists": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "ofp_ens_cells": (_i, [_vp, _u32p, _u32p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "ofp_ens_set": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "ofp_ens_get": (_i, [_vp, _vp, _vp, _vp, _sz, _vp]),
    "ofp_ens_rebuild": (_i, [_vp, _i, _vp]),
    "ofp_ens_forces": (_i, [_vp, _i, _vp, _vp]),
    "ofp_ens_pair_pass": (_i, [_vp, _i, _i, _vp]),
    "ofp_ens_run": (_i, [_vp, _i, C.c_double, _i, _i, _vp]),
    "ofp_ens_status": (_i, [_vp, C.POINTER(EnsState), _vp]),
    # test hook exported by the library but not declared in the public header
    "ofp_ens_test_masks_": (_i, [_vp, _sz, _i, C.POINTER(_sz)]),
}

_lib = None


def lib():
    """Load libofp_nn.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OFP_OK:
        raise OfpError(rc, lib().ofp_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
