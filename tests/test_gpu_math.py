"""The kernels' short correctly rounded operations of the Jacobi rotation equal the IEEE ones bit for
bit over every float of several exponent ranges (each range covers every significand, so the result
holds at every exponent where the short form is used): sqrt_ge1 (sqrt(theta^2 + 1)) against sqrtf
over [1, 4), [2^62, 2^64), [2^126, 2^128) and +inf; rcp_ge1 (1 / (|theta| + sqrt(theta^2 + 1)))
against 1.0f / d over [1, 2), [2^20, 2^21), [2^60, 2^61) and [2^125, 2^128) with +inf (the last
taking its IEEE branch). The kernels' SVD matches the oracle's bit for bit where Jacobi rotations are
skipped in some lanes of a wavefront."""
import ctypes
import os

import numpy as np
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


def _svd_inputs(n: int = 4096) -> np.ndarray:
    """Matrices whose Jacobi rotations are skipped (|a_pq| < 1e-30) in all, some or none of the
    rotations, cycling every 8 so that each wavefront holds all kinds: the identity; positive
    diagonals; diagonals with negative entries and signed zeros off the diagonal; off-diagonals of
    1e-32 (a_pq nonzero but under the threshold); one coupled pair (the other two planes skipped);
    near-identity and general shears."""
    rng = np.random.default_rng(11)
    out = np.empty((n, 3, 3), np.float32)
    for i in range(n):
        k = i % 8
        if k == 0:
            F = np.eye(3)
        elif k == 1:
            F = np.diag(rng.uniform(0.5, 1.5, 3))
        elif k == 2:
            F = np.diag(rng.uniform(0.5, 1.5, 3) * rng.choice([-1.0, 1.0], 3))
            F[0, 1], F[2, 0], F[1, 2] = -0.0, -0.0, 0.0
        elif k == 3:
            F = np.eye(3) + np.diag(rng.uniform(-0.2, 0.2, 3))
            F[0, 1], F[1, 2] = 1e-32, -1e-32
        elif k == 4:
            F = np.eye(3)
            F[:2, :2] += rng.normal(0.0, 0.2, (2, 2))
        elif k == 5:
            F = np.eye(3) + rng.normal(0.0, 0.03, (3, 3))
        else:
            F = np.eye(3) + rng.normal(0.0, 0.5, (3, 3))
        out[i] = F
    return out


def test_svd3_matches_oracle_with_skipped_rotations():
    """The kernels' svd3 (mpm_math.hpp; the Jacobi rotations computed by every lane, skipped lanes'
    values selected back) gives the oracle's U, s and V bit for bit, in wavefronts where every lane,
    some lanes or no lane skips a rotation."""
    import oracle_mpm as O
    print("[gpu] svd3 vs oracle, 4096 matrices with skipped rotations", flush=True)
    assert os.path.exists(LIB), "build the package first (it builds the check library)"
    L = ctypes.CDLL(LIB)
    fp = ctypes.POINTER(ctypes.c_float)
    L.run_svd3.argtypes = [fp, ctypes.c_uint32, fp, fp, fp]
    L.run_svd3.restype = ctypes.c_int
    F = _svd_inputs()
    n = len(F)
    Fc = np.ascontiguousarray(F.transpose(0, 2, 1)).reshape(-1)  # column-major per matrix
    U, s, V = np.zeros(9 * n, np.float32), np.zeros(3 * n, np.float32), np.zeros(9 * n, np.float32)
    ptr = lambda a: a.ctypes.data_as(fp)
    assert L.run_svd3(ptr(Fc), n, ptr(U), ptr(s), ptr(V)) == 0
    U, V, s = U.reshape(n, 3, 3).transpose(0, 2, 1), V.reshape(n, 3, 3).transpose(0, 2, 1), s.reshape(n, 3)
    skipped = 0
    for i in range(n):
        Uo, so, Vo = O.svd3(F[i])
        for name, got, want in (("U", U[i], Uo), ("s", s[i], so), ("V", V[i], Vo)):
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"matrix {i} (kind {i % 8}) {name}: {got} vs {want}"
        A = F[i].T.astype(np.float32) @ F[i].astype(np.float32)
        skipped += int(abs(A[0, 1]) < 1e-30 or abs(A[0, 2]) < 1e-30 or abs(A[1, 2]) < 1e-30)
    assert skipped >= n // 2
