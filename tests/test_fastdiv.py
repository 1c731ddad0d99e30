"""The fixed-divisor division the SDF kernels use (div_fixed: a correctly
rounded reciprocal, a product and two FMA residual corrections, IEEE division
outside [2^-100, FLT_MAX]) is bit-identical to IEEE division.  A C
restatement of the device function is compared with IEEE division on the
host, for every numerator significand of binades chosen to cover denormal,
tiny (down to the 2^-96 numerator and 2^-100 quotient bounds), ordinary and
overflowing quotients, and for the divisors the kernels use (grid cells, their
0.1 gradient steps) plus adversarial, out-of-range and random ones.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path_factory.mktemp("fastdiv") / "fastdiv_check")
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "cpp", "fastdiv_check.c"),
                    "-lm", "-lpthread"], check=True)
    return exe


def _f32(x):
    return float(np.float32(x))


def _divisors():
    cells = [0.0025, 0.005, 0.01, 0.32 / 64, 0.32 / 80, 0.32 / 96, 0.32 / 112, 0.32 / 128]
    out = [_f32(c) for c in cells] + [_f32(np.float32(0.1) * np.float32(c)) for c in cells]
    out += [struct.unpack("<f", struct.pack("<I", u))[0] for u in (0x3F7FFFFF, 0x3F800001, 0x3E3FFFFF, 0x40400000)]
    out += [2.0 ** -101, 2.0 ** 101, 1e-30, 1e30]        # outside the divisor range: IEEE division throughout
    rng = np.random.Generator(np.random.PCG64(11))
    bits = (rng.integers(110, 140, 24).astype(np.uint32) << 23) | rng.integers(0, 1 << 23, 24).astype(np.uint32)
    out += [float(x) for x in bits.view(np.float32)]
    return [float.hex(x) for x in out]


@pytest.mark.parametrize("binades", [(0, 2), (22, 33), (126, 128), (200, 202), (253, 255)])
def test_fixed_divisor_quotient_is_ieee(checker, binades):
    r = subprocess.run([checker, str(binades[0]), str(binades[1])] + _divisors(), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
