"""The kernels' short correctly rounded operations of the Jacobi rotation equal the IEEE ones bit for
bit over every float of several exponent ranges (each range covers every significand, so the result
holds at every exponent where the short form is used): sqrt_ge1 (sqrt(theta^2 + 1)) against sqrtf
over [1, 4), [2^62, 2^64), [2^126, 2^128) and +inf; rcp_ge1 (1 / (|theta| + sqrt(theta^2 + 1)))
against 1.0f / d over [1, 2), [2^20, 2^21), [2^60, 2^61) and [2^125, 2^128) with +inf (the last
taking its IEEE branch)."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "_build", "libgpu_math_check.so")

pytestmark = pytest.mark.gpu

CASES = [(0, 0x3F800000, 0x40800000), (0, 0x5E800000, 0x5F800000), (0, 0x7E800000, 0x7F800001),
         (1, 0x3F800000, 0x40000000), (1, 0x49800000, 0x4A000000), (1, 0x5D800000, 0x5E000000),
         (1, 0x7E000000, 0x7F800001)]


@pytest.mark.parametrize("which,lo,hi", CASES)
def test_short_math_matches_ieee(which, lo, hi):
    name = ("sqrt_ge1", "rcp_ge1")[which]
    print(f"[gpu] {name} over bits {lo:#x}..{hi:#x}", flush=True)
    assert os.path.exists(LIB), "build the package first (it builds the check library)"
    L = ctypes.CDLL(LIB)
    L.check_short_math.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ctypes.c_uint32)]
    L.check_short_math.restype = ctypes.c_int
    bad, first = ctypes.c_uint64(), ctypes.c_uint32()
    assert L.check_short_math(which, lo, hi - lo, ctypes.byref(bad), ctypes.byref(first)) == 0
    assert bad.value == 0, f"{name}: {bad.value} mismatches, first at bits {first.value:#x}"
