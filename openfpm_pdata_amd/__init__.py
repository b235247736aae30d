"""openfpm_pdata_amd: the vector_dist particle neighbour-search path of
openfpm_pdata on one MI355X (gfx950), as hand-written HIP kernels behind a C ABI
(libofp_nn.so, include/ofp_nn.h), a C++ facade (include/openfpm_amd/) and this
Python host layer."""
from . import _lib  # noqa: F401

__all__ = ["_lib", "vector_dist", "workloads"]
