"""The kick kernels divide by 6 and by pi (real(4)) as
q0 = a*y, q1 = fma(fma(-q0, b, a), y, q0) with y = RN(1/b), taking the
IEEE division only for |a| outside [1e-30, 1e30).  This checks on the CPU
that the two agree bit for bit: every mantissa at the ends and the middle of
that range (inside it every step scales exactly with the exponent), or, with
CUBE_EXHAUSTIVE=1, every float of the range (about a minute)."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PI_F = struct.unpack("<f", struct.pack("<f", np.float32(np.pi)))[0]


def _bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("div") / "divide_check.so")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-shared", "-fPIC",
                    os.path.join(HERE, "divide_check.c"), "-o", so, "-lm"], check=True)
    lib = ctypes.CDLL(so)
    lib.count_bad.argtypes = [ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    lib.count_bad.restype = ctypes.c_int64
    return lib


@pytest.mark.parametrize("b", [6.0, PI_F])
def test_fixed_divisor_form_is_exact(checker, b):
    lo, hi = _bits(1e-30), _bits(1e30)
    if os.environ.get("CUBE_EXHAUSTIVE") == "1":
        spans = [(lo, hi)]
    else:
        e0 = lo & ~0x7FFFFF
        spans = [(lo, e0 + (2 << 23)), (_bits(1.0), _bits(4.0)), (hi - (2 << 23), hi)]
    for a, z in spans:
        assert checker.count_bad(b, a, z, 0) == 0, (b, a, z)


def test_checker_sees_a_wrong_form(checker):
    # without the correction step the product a*RN(1/b) is often off by an ulp
    lo = _bits(1.0)
    assert checker.count_bad(6.0, lo, lo + (1 << 20), 1) > 1000
    assert checker.count_bad(PI_F, lo, lo + (1 << 20), 1) > 1000
