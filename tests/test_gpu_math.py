"""The kernels' short square root (sqrt_ge1, used for the Jacobi rotation's sqrt(theta^2 + 1)) equals
sqrtf bit for bit over every float in [1, 4) (every significand at both exponent parities),
[2^62, 2^64), [2^126, 2^128) and at +inf."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "_build", "libgpu_math_check.so")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lo,hi", [(0x3F800000, 0x40800000), (0x5E800000, 0x5F800000), (0x7E800000, 0x7F800001)])
def test_sqrt_ge1_matches_sqrtf(lo, hi):
    print(f"[gpu] sqrt_ge1 over bits {lo:#x}..{hi:#x}", flush=True)
    assert os.path.exists(LIB), "build the package first (it builds the check library)"
    L = ctypes.CDLL(LIB)
    L.check_sqrt_ge1.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.POINTER(ctypes.c_uint32)]
    L.check_sqrt_ge1.restype = ctypes.c_int
    bad, first = ctypes.c_uint64(), ctypes.c_uint32()
    assert L.check_sqrt_ge1(lo, hi - lo, ctypes.byref(bad), ctypes.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} mismatches, first at bits {first.value:#x}"
