"""The C++ classes of include/tressfx_amd.hpp driven through the reference's
per-frame order by a compiled program (tests/cpp/shim_test.cpp), bit-identical
to the same frames through the C-ABI."""
import os
import subprocess

import pytest

from tressfx_amd import build

pytestmark = pytest.mark.gpu


def test_cpp_shim_program(gpu):
    exe = build.CPP_TEST_BIN
    assert os.path.exists(exe), "build the C++ shim test first (python -m tressfx_amd.build)"
    print("[cpp] tests/cpp/shim_test", flush=True)
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    res = subprocess.run([exe, golden], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert res.returncode == 0 and "shim_test OK" in res.stdout, res.stdout
