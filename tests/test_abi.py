"""C-ABI checks that need no GPU: the library loads, exports every function
include/ofp_nn.h declares, and its host-only geometry entry points equal the
oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as ora
from openfpm_pdata_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ofp_nn.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ofp_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    lib = L.lib()
    for n in names:
        assert hasattr(lib, n), n
    # and the binding knows every one of them
    assert set(names) <= set(L.SIGNATURES), set(names) - set(L.SIGNATURES)


def test_io_library_exports_every_declared_symbol():
    import ctypes

    from openfpm_pdata_amd import h5io

    src = open(os.path.join(ROOT, "include", "ofp_io.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(ofp_[a-z0-9_]+)\s*\(", src)))
    assert len(names) == 4
    lib = h5io.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert isinstance(lib, ctypes.CDLL)


def test_abi_version_and_errors():
    lib = L.lib()
    assert lib.ofp_abi_version() == 1
    rc = lib.ofp_shift_vectors(7, None, None, None)
    assert rc == L.OFP_E_ARG
    assert b"null" in lib.ofp_last_error()


def test_ensemble_argument_errors():
    """ofp_ens_create rejects bad descriptions before touching the device."""
    lib = L.lib()
    h = C.c_void_p()

    def desc(n=100, rc=0.1, side=1.0, sigma=0.04, mass=1.0):
        return (L.SysDesc * 1)(L.SysDesc((C.c_double * 3)(0, 0, 0), (C.c_double * 3)(side, side, side), rc, sigma,
                                         1.0, mass, n))

    assert lib.ofp_ens_create(None, L.OFP_F32, 1, desc(), None) == L.OFP_E_ARG
    assert lib.ofp_ens_create(C.byref(h), 5, 1, desc(), None) == L.OFP_E_ARG
    assert lib.ofp_ens_create(C.byref(h), L.OFP_F32, 0, desc(), None) == L.OFP_E_ARG
    # the box must hold twice the Verlet radius
    assert lib.ofp_ens_create(C.byref(h), L.OFP_F32, 1, desc(rc=0.5), None) == L.OFP_E_LIMIT
    assert b"twice the Verlet radius" in lib.ofp_last_error()
    assert lib.ofp_ens_create(C.byref(h), L.OFP_F32, 1, desc(mass=0.0), None) == L.OFP_E_ARG
    bad = L.EnsOpts(0.1, 2, 12, 32, 0, 1.25)
    assert lib.ofp_ens_create(C.byref(h), L.OFP_F32, 1, desc(), C.byref(bad)) == L.OFP_E_ARG
    bad = L.EnsOpts(0.1, 4, 16, 32, 0, 1.25)
    assert lib.ofp_ens_create(C.byref(h), L.OFP_F32, 1, desc(), C.byref(bad)) == L.OFP_E_LIMIT
    assert lib.ofp_ens_destroy(None) == L.OFP_OK


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_geometry_matches_oracle(dt):
    import torch

    from openfpm_pdata_amd import vector_dist as vdm

    tdt = {np.float32: torch.float32, np.float64: torch.float64}[dt]
    for lo, hi, g, bc in (((0, 0, 0), (1, 1, 1), 0.05, (1, 1, 1)), ((-0.5, 0, 2), (0.5, 2, 3), 0.11, (1, 0, 1)),
                          ((0, 0, 0), (1, 1, 1), 0.0161, (0, 0, 0))):
        np.testing.assert_array_equal(vdm.shift_vectors(tdt, lo, hi), ora.shift_vectors(lo, hi, dt))
        a = vdm.local_ghost_boxes(tdt, lo, hi, [-g] * 3, [g] * 3, bc)
        b = ora.local_ghost_boxes(lo, hi, [-g] * 3, [g] * 3, bc, dt)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
        for rc, f in ((g, float(np.float32(1.013))), (g * 1.1, 1.013)):
            blo, bhi, div = vdm.cl_param(tdt, lo, hi, rc, [-g] * 3, [g] * 3, f)
            geom = ora.CellGeom(lo, hi, g, rc, f, dt)
            np.testing.assert_array_equal(blo, geom.lo)
            np.testing.assert_array_equal(bhi, geom.hi)
            np.testing.assert_array_equal(div, geom.div)


def test_cl_param_rejects_bad_input():
    lo = np.zeros(3)
    hi = np.ones(3)
    out = np.zeros(3)
    div = np.zeros(3, np.uint32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = L.lib().ofp_cl_param(L.OFP_F64, p(lo), p(hi), -1.0, p(lo), p(lo), 1.0, p(out), p(out), p(div))
    assert rc == L.OFP_E_ARG
