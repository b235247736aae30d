"""The C++ facade exercised as user code: tests/cpp/test_vector_dist_gpu.cpp
is user code in the style of openfpm_pdata GPU programs (ghost_get, cell lists,
toKernel() views, GET_PARTICLE(_SORT), NN iterators with get/get_sort,
getVerletGPU, getVerletSym, ghost_put, save / load), compiled with hipcc against
include/openfpm_amd/vector_dist.hpp (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_vector_dist_gpu")


def test_facade_binary_built():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_facade_user_code_on_gpu():
    golden = os.path.join(ROOT, "tests", "golden", "vector_dist_24.h5")
    r = subprocess.run([BIN, golden], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
