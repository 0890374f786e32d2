"""Pins of the oracle's building blocks against mathematics and independent references (CPU).

- exp/log: float64 numpy within a few ulp over wide ranges and special values.
- morton3: brute-force bit interleave over every 10-bit input (mpm.js:70-78).
- toFixed: closed forms of round(clamp(x * 2^22)) including ties to even (mpm.js:66).
- bspline: partition of unity and zero first moment (docs/GUIDE.md:265-293 relies on both).
- svd3: reconstruction, orthogonality, rotation signs, ordering, and singular values against
  numpy.linalg.svd (linalg.js:46-69 has no numeric test in the reference).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_mpm as O


def ulp_err(got: np.ndarray, want64: np.ndarray) -> np.ndarray:
    want32 = want64.astype(np.float32)
    spacing = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want64) / spacing


def test_expf_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-87, 88, 20000), rng.uniform(-2, 2, 20000), rng.uniform(-1e-3, 1e-3, 2000),
                         np.array([0.0, -0.0, 1.0, -1.0, 0.5, 88.0, -87.0, 1e-30, -1e-30])]).astype(np.float32)
    got = np.array([O.expf(float(x)) for x in xs], np.float32)
    err = ulp_err(got, np.exp(xs.astype(np.float64)))
    assert err.max() <= 3.0, (err.max(), xs[np.argmax(err)])
    assert O.expf(0.0) == 1.0
    assert O.expf(200.0) == np.inf
    assert O.expf(-200.0) == 0.0
    assert np.isnan(O.expf(float("nan")))


def test_logf_accuracy():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), rng.uniform(0.5, 2.0, 20000),
                         np.array([1.0, 2.0, 0.5, 1e-4, 1e-38, 1e-44, 3.4e38])]).astype(np.float32)
    xs = xs[xs > 0]
    got = np.array([O.logf(float(x)) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64))
    err = ulp_err(got, want)
    # near x = 1 the result is tiny; measure against an absolute floor of one ulp of 1e-7
    err = np.where(np.abs(want) < 1e-6, np.abs(got - want) / 1e-13, err)
    assert err.max() <= 3.0, (err.max(), xs[np.argmax(err)])
    assert O.logf(1.0) == 0.0
    assert O.logf(0.0) == -np.inf
    assert np.isnan(O.logf(-1.0))


def test_rsqrt_1to2_every_float_in_range():
    """inverseSqrt on [1, 2] (rsqrt_1to2, the Jacobi rotation's c = inverseSqrt(t^2 + 1)): every one of
    the 2^23 + 1 floats in [1, 2] is within 1 ulp of the float64 1/sqrt (WGSL allows 2 ulp), and
    more than 80% of them (84.8%) are the correctly rounded value (a plausible slip, e.g. a wrong coefficient
    sign or a dropped Newton step, breaks one or both)."""
    x = (np.arange(2 ** 23 + 1, dtype=np.uint32) + np.uint32(0x3F800000)).view(np.float32)
    assert x[0] == 1.0 and x[-1] == 2.0
    y = O.rsqrt_1to2(x)
    ref = 1.0 / np.sqrt(x.astype(np.float64))
    err = np.abs(y.astype(np.float64) - ref) / np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert err.max() < 1.0, err.max()
    assert (y == ref.astype(np.float32)).mean() > 0.8
    assert O.rsqrt_1to2(np.array([np.nan], np.float32))[0] != O.rsqrt_1to2(np.array([np.nan], np.float32))[0]


def test_morton3_brute_force():
    def interleave(x, y, z):
        k = 0
        for b in range(10):
            k |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        return k

    for v in range(1024):
        assert O.lib().or_morton3(v, 0, 0) == interleave(v, 0, 0)
        assert O.lib().or_morton3(0, v, 0) == interleave(0, v, 0)
        assert O.lib().or_morton3(0, 0, v) == interleave(0, 0, v)
    rng = np.random.default_rng(2)
    for x, y, z in rng.integers(0, 1024, (2000, 3)):
        assert O.lib().or_morton3(int(x), int(y), int(z)) == interleave(int(x), int(y), int(z))


def test_to_fixed_closed_forms():
    tf = O.lib().or_to_fixed
    FP = 2.0 ** 22
    assert tf(1.0) == 1 << 22
    assert tf(-1.0) == -(1 << 22)
    assert tf(0.5 / FP) == 0          # tie to even
    assert tf(1.5 / FP) == 2
    assert tf(2.5 / FP) == 2
    assert tf(-2.5 / FP) == -2
    assert tf(1e30) == 2000000000     # clamp
    assert tf(-1e30) == -2000000000
    rng = np.random.default_rng(3)
    for x in rng.uniform(-100, 100, 2000).astype(np.float32):
        assert tf(float(x)) == int(np.rint(np.float32(x) * np.float32(FP)))


def test_bspline_partition_of_unity():
    rng = np.random.default_rng(4)
    w = np.zeros(9, np.float32)
    for fx in rng.uniform(0.5, 1.5, (500, 3)).astype(np.float32):
        fx = np.ascontiguousarray(fx)
        O.lib().or_bspline(O._p(fx), O._p(w))
        W = w.reshape(3, 3)
        for ax in range(3):
            assert abs(W[:, ax].sum() - 1.0) < 1e-6
            moment = sum(W[k, ax] * (k - fx[ax]) for k in range(3))
            assert abs(moment) < 1e-6
            assert np.all(W[:, ax] >= 0)


def _rand_F(rng, n, scale):
    return np.eye(3)[None] + rng.uniform(-scale, scale, (n, 3, 3))


# F = I + uniform(-scale, scale): the method (Jacobi on F^T F in float32, 4 sweeps) loses accuracy as F
# departs from a rotation, so the bound grows with the deformation; a wrong sign, index or ordering
# gives O(1) errors and fails every bound.
@pytest.mark.parametrize("scale,tol", [(0.05, 1e-5), (0.3, 2e-5), (0.8, 1e-4)])
def test_svd3_reconstruction_and_numpy(scale, tol):
    rng = np.random.default_rng(int(scale * 100))
    Fs = _rand_F(rng, 3000, scale).astype(np.float32)
    worst_rec, worst_sv = 0.0, 0.0
    for F in Fs:
        U, s, V = O.svd3(F)
        rec = U @ np.diag(s) @ V.T
        worst_rec = max(worst_rec, np.abs(rec - F).max() / max(1.0, np.abs(F).max()))
        assert abs(np.linalg.det(V) - 1) < 1e-4
        assert abs(np.linalg.det(U) - 1) < 1e-4
        assert np.abs(U.T @ U - np.eye(3)).max() < 1e-4
        assert np.abs(V.T @ V - np.eye(3)).max() < 1e-4
        assert s[0] >= s[1] >= abs(s[2]) - 1e-6
        ref = np.linalg.svd(F.astype(np.float64), compute_uv=False)
        sref = ref.copy()
        if np.linalg.det(F.astype(np.float64)) < 0:
            sref[2] = -sref[2]
        worst_sv = max(worst_sv, np.abs(s - sref).max() / max(1e-3, ref[0]))
    assert worst_rec < tol, worst_rec
    assert worst_sv < tol, worst_sv


def test_svd3_special_cases():
    # identity, rotation, diagonal with repeated values, reflection
    U, s, V = O.svd3(np.eye(3, dtype=np.float32))
    assert np.allclose(s, 1) and np.allclose(U @ V.T, np.eye(3), atol=1e-6)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]], np.float32)
    U, s, V = O.svd3(R)
    assert np.allclose(s, 1, atol=1e-6) and np.allclose(U @ V.T, R, atol=1e-5)
    D = np.diag([2.0, 2.0, 0.5]).astype(np.float32)
    U, s, V = O.svd3(D)
    assert np.allclose(s, [2, 2, 0.5], atol=1e-6)
    M = np.diag([1.0, 1.0, -1.0]).astype(np.float32)
    U, s, V = O.svd3(M)
    assert np.allclose(np.abs(s), 1, atol=1e-6) and s[2] < 0  # det < 0 carried by the last value
    assert np.allclose(U @ np.diag(s) @ V.T, M, atol=1e-5)
