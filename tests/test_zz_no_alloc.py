"""mpm_scan_u32_noalloc (the scan with its partial sums in the library's fixed device array) against
mpm_scan_u32 (Scan.run, src/prims/scan.js:292-306) through the C ABI: bit-identical outputs on two
calls in a row, the second with a smaller input, in and out of place; the size limit is refused
before any work. Device buffers are allocated here by the test, never by the entry point."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return hip


def _scan_host(ctx, a: np.ndarray) -> np.ndarray:
    from shaderintel_amd.mpm import load_library, _ptr, _check
    out = np.empty_like(a)
    _check(load_library().mpm_scan_u32(ctx._h, _ptr(a), _ptr(out), a.size), "mpm_scan_u32")
    return out


def _scan_noalloc(ctx, hip, a: np.ndarray, in_place: bool) -> np.ndarray:
    from shaderintel_amd.mpm import load_library, _ptr, _check
    L = load_library()
    L.mpm_scan_u32_noalloc.restype = ctypes.c_int32
    L.mpm_scan_u32_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                       ctypes.c_void_p]
    nbytes = a.size * 4
    d_in, d_out = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_in), nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(d_out), nbytes) == 0
    try:
        assert hip.hipMemcpy(d_in, _ptr(a), nbytes, H2D) == 0
        dst = d_in if in_place else d_out
        _check(L.mpm_scan_u32_noalloc(ctx._h, d_in, dst, a.size, None), "mpm_scan_u32_noalloc")
        assert hip.hipStreamSynchronize(None) == 0
        out = np.empty_like(a)
        assert hip.hipMemcpy(_ptr(out), dst, nbytes, D2H) == 0
        if not in_place:  # the input is not written
            back = np.empty_like(a)
            assert hip.hipMemcpy(_ptr(back), d_in, nbytes, D2H) == 0
            assert np.array_equal(back, a)
        return out
    finally:
        hip.hipFree(d_in)
        hip.hipFree(d_out)


@pytest.mark.parametrize("in_place", [False, True])
def test_noalloc_scan_matches_scan(ctx, in_place):
    hip = _hip()
    rng = np.random.default_rng(11)
    # 20,000,011 elements: 4,883 tiles, so the partial sums are scanned recursively; then 4,097 (two
    # tiles, offsets added up per tile) and 1,000 (one tile), each after a larger call
    for n in (20_000_011, 4_097, 1_000):
        print(f"[gpu] no-alloc scan n={n} in_place={in_place}", flush=True)
        a = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)  # wraps like the reference's u32
        want = _scan_host(ctx, a)
        got = _scan_noalloc(ctx, hip, a, in_place)
        assert np.array_equal(got, want), f"n={n}: {int((got != want).sum())} words differ"


def test_noalloc_scan_refuses_beyond_maximum(ctx):
    from shaderintel_amd.mpm import load_library
    L = load_library()
    L.mpm_scan_u32_noalloc.restype = ctypes.c_int32
    L.mpm_scan_u32_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                       ctypes.c_void_p]
    # non-null placeholders: the size check comes before anything touches them
    st = L.mpm_scan_u32_noalloc(ctx._h, ctypes.c_void_p(16), ctypes.c_void_p(16), (1 << 28) + 1, None)
    assert st == 1
    assert b"exceeds the maximum" in L.mpm_last_error_message()
    assert L.mpm_scan_u32_noalloc(ctx._h, None, None, 0, None) == 0  # n = 0: nothing to do
