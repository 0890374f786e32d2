"""Pins the oracle's P2G word for word, and its G2P sums within the spread of faithful readings,
against tests/literal.py: numpy restatements written node by node from the WGSL of
src/physics/mpm.js:318-349 and :483-497, sharing no code with oracle/mpm_oracle.cpp. The fused
multiply-add emulation they rest on is itself checked against exact rational arithmetic. CPU only."""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

import literal as L
import oracle_mpm as O
from helpers import randomize_fast

DX = 1 / 64


def _round_f32(x: Fraction) -> np.float32:
    """Round an exact rational to binary32, nearest even, by integer arithmetic (no float on the way)."""
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()  # 2^e <= a < 2^(e+2)
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)  # subnormals share the smallest exponent
    q = a / Fraction(2) ** (e - 23)
    n = round(q)  # Fraction rounds half to even
    return np.float32(sign * float(Fraction(n) * Fraction(2) ** (e - 23)))


def test_fma32_is_correctly_rounded():
    """fma32 against the exactly rounded a*b + c on random triples, cancellations, exact ties and
    subnormal results."""
    rng = np.random.default_rng(7)
    a = (rng.standard_normal(3000) * 10.0 ** rng.integers(-6, 6, 3000)).astype(np.float32)
    b = (rng.standard_normal(3000) * 10.0 ** rng.integers(-6, 6, 3000)).astype(np.float32)
    c = (rng.standard_normal(3000) * 10.0 ** rng.integers(-6, 6, 3000)).astype(np.float32)
    c[:1000] = -(a[:1000] * b[:1000])  # heavy cancellation: the result is the product's rounding error
    one_ulp = np.float32(2.0 ** -23)
    extra = [
        (np.float32(1.0) + one_ulp, np.float32(1.0) + one_ulp, np.float32(-1.0)),  # 2^-22 + 2^-46, exact
        (np.float32(1.0) + one_ulp, np.float32(1.0) - one_ulp / 2, np.float32(0.0)),  # 1 + 2^-24 - 2^-47
        (np.float32(3.0), np.float32(1.0) + one_ulp, np.float32(2.0 ** -22)),  # a tie broken by the product
        (np.float32(1e-30), np.float32(1e-10), np.float32(0.0)),  # subnormal result
        (np.float32(0.5), np.float32(2.0 ** -149), np.float32(0.0)),  # half the smallest subnormal: ties to 0
    ]
    a = np.concatenate([a, np.array([t[0] for t in extra], np.float32)])
    b = np.concatenate([b, np.array([t[1] for t in extra], np.float32)])
    c = np.concatenate([c, np.array([t[2] for t in extra], np.float32)])
    got = L.fma32(a, b, c)
    want = np.array([_round_f32(Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z)))
                     for x, y, z in zip(a, b, c)], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the plain two-rounding product-sum differs somewhere, so the test can see a missing fusion
    assert not np.array_equal((a * b + c).view(np.uint32), want.view(np.uint32))


def _dam64(randomize: bool = True):
    dims = [64, 64, 64]
    pts = O.sample_shape(0, [3 * DX, 3 * DX, 3 * DX, 11 * DX, 35 * DX, 35 * DX], DX / 2, 1.0, 1)
    rec = O.pack_particles(pts, [0, 0, 0], DX, dims, 3.0, 0, [0, 0, 0], (1 / 480) / DX)
    if randomize:
        randomize_fast(rec, seed=3)
    tbl = np.zeros((8, 12), np.float32)
    tbl[0] = O.material_row({"kind": 0, "density": 1.0, "stiffness": 1.0, "viscosity": 0.0}, 2)
    return dims, rec, tbl


def _oracle_p2g(dims, rec, tbl, ppc=2):
    sc = O.Scene(dims, len(rec), ppc, 4)
    sc.set_materials(tbl)
    sc.write(rec)
    sc.set_count(len(rec))
    sc.p2g()
    return sc


def test_p2g_words_equal_per_node_reading_dam64():
    """The C1 dam break (65,536 water particles) with random v, C and tau (physics.test.js:9-17's
    ranges): every one of the oracle's i32 grid words equals the per-node reading."""
    dims, rec, tbl = _dam64()
    got = _oracle_p2g(dims, rec, tbl).grid()
    want = L.p2g(rec, tbl, 0.5, dims)
    assert (want != 0).sum() > 50000
    assert np.array_equal(got, want), f"{int((got != want).sum())} words differ"


def test_p2g_words_equal_per_node_reading_materials_and_clamp():
    """Four materials with sheared F and full tau (the golden scene's construction), plus particles
    fast enough that toFixed's clamp acts on some contributions, and ppc 3 (vol4 = 4/27)."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden
    rec, tbl, _ = make_golden.build()
    dims = make_golden.DIMS
    fast = rec[:40].copy()
    fast[:, 4:7] = np.float32(3.0e4) * np.sign(fast[:, 4:7] + np.float32(1e-3))  # m v up to 6000: clamps
    fast[:, 8:11] *= np.float32(1.0e5)
    rec = np.concatenate([rec, fast])
    for ppc, vol4 in ((2, 0.5), (3, np.float32(4.0 / 27.0))):
        got = _oracle_p2g(dims, rec, tbl, ppc).grid()
        stats: dict = {}
        want = L.p2g(rec, tbl, vol4, dims, stats)
        assert np.array_equal(got, want), f"ppc {ppc}: {int((got != want).sum())} words differ"
        assert stats["clamped"] > 100  # the fast particles do reach toFixed's clamp


def test_p2g_per_node_reading_sees_a_regrouped_sum():
    """The check has teeth: the regrouped momentum sum ((u + a A0) + b A1) + c A2 with
    u = mv - affine fx, the same terms in another order, changes words of the randomized dam64
    (restated here in numpy, separably as a kernel might)."""
    dims, rec, tbl = _dam64()
    X, Y, _ = dims
    xp = rec[:, 0:3]
    base = np.floor(xp - np.float32(0.5)).astype(np.int64)
    fx = xp - base.astype(np.float32)
    w = L.bspline(fx)
    mass = tbl[0, 1]
    tau, C = L._tau_C(rec)
    affine = np.float32(-0.5) * tau + mass * C
    mv = mass * rec[:, 4:7]
    u = mv - L.mat_vec(affine, fx)
    acc = np.zeros((np.prod(dims), 4), np.int64)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                U = u.copy()
                for k, o in ((0, a), (1, b), (2, c)):
                    if o:
                        U = U + np.float32(o) * affine[:, k, :]
                wt = (w[a][:, 0] * w[b][:, 1]) * w[c][:, 2]
                q = np.concatenate([L.to_fixed(wt[:, None] * U), L.to_fixed(wt * mass)[:, None]], axis=1)
                g = base + np.array([a, b, c])
                np.add.at(acc, g[:, 0] + X * (g[:, 1] + Y * g[:, 2]), q)
    regrouped = acc.astype(np.int32)
    want = L.p2g(rec, tbl, 0.5, dims)
    assert np.array_equal(regrouped[:, 3], want[:, 3])  # mass words do not depend on it
    assert (regrouped[:, :3] != want[:, :3]).sum() > 100


def test_g2p_sums_within_faithful_readings():
    """G2P's v and C: the oracle sums per row of three nodes (S = sum_c wz gv, T = sum_c wz dz gv);
    the reference writes v += wt gv, B += outer(wt gv, dpos) node by node and leaves the fusion of
    those accumulations to the compiler. The oracle's result must lie within the spread of the two
    node-by-node readings (fused and unfused) after one P2G and grid update of the randomized dam64."""
    dims, rec, tbl = _dam64()
    sc = _oracle_p2g(dims, rec, tbl)
    g = np.array([0.0, np.float32(-9.81 * (1 / 480) ** 2 / DX), 0.0], np.float32)
    sc.grid_update(g, 0.4)
    gv = sc.grid().view(np.float32)
    X, Y, _ = dims

    def gv_at(x, y, z):
        return gv[x + X * (y + Y * z), :3]

    sc.g2p(100.0)  # vmax out of reach: v and C as gathered
    out = sc.read(len(rec))
    v_or = out[:, 4:7].astype(np.float64)
    C_or = np.stack([out[:, 8:11], out[:, 12:15], out[:, 16:19]], axis=1).astype(np.float64)
    v_f, C_f = L.g2p_velocity(rec, gv_at, fused=True)
    v_u, C_u = L.g2p_velocity(rec, gv_at, fused=False)
    # a priori bound on the rounding of any evaluation order of a 27-term sum (Higham, "Accuracy
    # and Stability of Numerical Algorithms", 2nd ed., eq. 4.4): |err| <= gamma_n sum |term| with
    # gamma_n = n u / (1 - n u); two orders differ by at most twice that (n = 27 terms, plus the
    # weight products: n = 30, u = 2^-24)
    xp = rec[:, 0:3]
    base = np.floor(xp - np.float32(0.5)).astype(np.int64)
    fx = (xp - base.astype(np.float32)).astype(np.float64)
    w = [0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1) ** 2, 0.5 * (fx - 0.5) ** 2]
    sv = np.zeros((len(rec), 3))
    sB = np.zeros((len(rec), 3, 3))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                wt = w[a][:, 0] * w[b][:, 1] * w[c][:, 2]
                dpos = np.abs(np.array([a, b, c]) - fx)
                gabs = np.abs(gv_at(base[:, 0] + a, base[:, 1] + b, base[:, 2] + c).astype(np.float64))
                sv += wt[:, None] * gabs
                sB += 4 * wt[:, None, None] * dpos[:, :, None] * gabs[:, None, :]
    u = 2.0 ** -24
    gamma = 30 * u / (1 - 30 * u)
    for name, ours, bound in (("v", v_or, sv), ("C", C_or, sB)):
        for lit in ((v_f, v_u) if name == "v" else (C_f, C_u)):
            assert np.all(np.abs(ours - lit) <= 2 * gamma * bound + 1e-30), name
    spread_v = np.abs(v_f.astype(np.float64) - v_u).max()
    spread_C = np.abs(C_f.astype(np.float64) - C_u).max()
    dv = max(np.abs(v_or - v_f).max(), np.abs(v_or - v_u).max())
    dC = max(np.abs(C_or - C_f).max(), np.abs(C_or - C_u).max())
    print(f"G2P envelope: spread of readings v {spread_v:.3g} C {spread_C:.3g}; oracle off by v {dv:.3g} C {dC:.3g}; "
          f"|v| max {np.abs(v_f).max():.3g}, |C| max {np.abs(C_f).max():.3g}")
    assert spread_v > 0 and spread_C > 0  # the two readings do differ: the bound is not vacuous
    assert dv <= 1e-5 * np.abs(v_f).max() and dC <= 1e-5 * np.abs(C_f).max()
