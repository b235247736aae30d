"""ofp_reduce_noalloc (allocates nothing, result in device memory) against
ofp_reduce on the same inputs: the same kernels and grid, so the results are
bit-identical.  Two calls in a row, the second with a smaller input (the
per-GPU partials array keeps the first call's contents), for every operator,
both dtypes, with and without the compensation term, and n = 0."""
import ctypes as C

import numpy as np
import pytest
import torch

from openfpm_pdata_amd import _lib as L
from openfpm_pdata_amd import vector_dist as vdm

pytestmark = pytest.mark.gpu

OPS = (L.OFP_RED_SUM, L.OFP_RED_MAX, L.OFP_RED_MIN)


def _both(x, comp, n, op, out):
    dt = L.OFP_F32 if x.dtype == torch.float32 else L.OFP_F64
    ref = C.c_double(0.0)
    L.call("ofp_reduce", dt, x.data_ptr(), comp.data_ptr() if comp is not None else None, n, op, C.byref(ref),
           vdm._stream())
    out.fill_(np.nan)
    L.call("ofp_reduce_noalloc", dt, x.data_ptr(), comp.data_ptr() if comp is not None else None, n, op,
           out.data_ptr(), vdm._stream())
    torch.cuda.synchronize()
    return ref.value, float(out.item())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("with_comp", [False, True])
def test_noalloc_matches_reduce_two_calls(dtype, with_comp):
    g = torch.Generator(device="cuda").manual_seed(7)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    for n in (3_000_017, 40_961, 1):  # second and third calls smaller than the first
        x = (torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1).to(dtype)
        comp = (torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 1e-7).to(dtype) if with_comp else None
        for op in OPS:
            ref, got = _both(x, comp, n, op, out)
            print(f"noalloc reduce dtype={dtype} comp={with_comp} n={n} op={op}: {got!r} vs {ref!r}", flush=True)
            assert np.float64(got).tobytes() == np.float64(ref).tobytes(), (n, op, got, ref)
        # the FP64 sum against numpy (a check independent of both entry points)
        xs = x.double().cpu().numpy() + (comp.double().cpu().numpy() if comp is not None else 0.0)
        _, got = _both(x, comp, n, L.OFP_RED_SUM, out)
        assert abs(got - xs.sum()) <= 1e-12 * max(1.0, np.abs(xs).sum())


def test_noalloc_empty_and_limits():
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    x = torch.ones(8, dtype=torch.float32, device="cuda")
    for op, ident in ((L.OFP_RED_SUM, 0.0), (L.OFP_RED_MAX, -np.inf), (L.OFP_RED_MIN, np.inf)):
        out.fill_(np.nan)
        L.call("ofp_reduce_noalloc", L.OFP_F32, x.data_ptr(), None, 0, op, out.data_ptr(), vdm._stream())
        torch.cuda.synchronize()
        assert float(out.item()) == ident
    with pytest.raises(L.OfpError):  # beyond the stated maximum: rejected before any launch
        L.call("ofp_reduce_noalloc", L.OFP_F32, x.data_ptr(), None, (1 << 40) + 1, L.OFP_RED_SUM, out.data_ptr(),
               vdm._stream())
    with pytest.raises(L.OfpError):
        L.call("ofp_reduce_noalloc", L.OFP_F32, x.data_ptr(), None, 8, L.OFP_RED_SUM, None, vdm._stream())
