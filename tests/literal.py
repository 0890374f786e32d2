"""Independent numpy restatements of the reference's MLS-MPM transfers, written node by node as the
WGSL in src/physics/mpm.js states them. They share no code with the oracle (oracle/mpm_oracle.cpp)
and are used to pin it: P2G word for word, G2P within the spread of faithful float readings.

Float32 arithmetic is numpy's (IEEE-754 binary32, round to nearest even). A fused multiply-add is
emulated exactly by fma32: the product of two binary32 values is exact in binary64, the sum is
rounded to odd in binary64 (the TwoSum error decides the sticky bit), and rounding that to binary32
gives the correctly rounded fused result (round-to-odd with 53 >= 24 + 2 bits leaves no double-
rounding error; Boldo and Melquiond, "Emulation of FMA and correctly rounded sums: proved
algorithms using rounding to odd", IEEE TC 2008).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
FP = F32(4194304.0)  # 2^22 (mpm.js:64)


def fma32(a, b, c) -> np.ndarray:
    """Correctly rounded binary32 a*b + c, elementwise."""
    a64 = np.asarray(a, np.float32).astype(np.float64)
    b64 = np.asarray(b, np.float32).astype(np.float64)
    c64 = np.asarray(c, np.float32).astype(np.float64)
    p = a64 * b64  # exact: 24 + 24 significant bits
    s = p + c64
    bb = s - p
    e = (p - (s - bb)) + (c64 - bb)  # TwoSum: p + c = s + e exactly
    even = (s.view(np.int64) & 1) == 0
    fix = (e != 0) & even & np.isfinite(s)
    toward = np.where(e > 0, np.inf, -np.inf)
    s = np.where(fix, np.nextafter(s, toward), s)
    return s.astype(np.float32)


def to_fixed(x) -> np.ndarray:
    """toFixed (mpm.js:64-66): i32(round(clamp(x * 2^22, -2e9, 2e9))), round half to even."""
    y = np.asarray(x, np.float32) * FP
    y = np.minimum(np.maximum(y, F32(-2.0e9)), F32(2.0e9))
    return np.rint(y).astype(np.int64)


def bspline(fx: np.ndarray) -> list[np.ndarray]:
    """Quadratic B-spline weights (mpm.js:67-72), each (n, 3) float32."""
    t0 = F32(1.5) - fx
    t1 = fx - F32(1.0)
    t2 = fx - F32(0.5)
    return [(F32(0.5) * t0) * t0, F32(0.75) - t1 * t1, (F32(0.5) * t2) * t2]


def mat_vec(M: np.ndarray, v: np.ndarray) -> np.ndarray:
    """M * v for WGSL mat3x3f (M[..., col, row]) as ((M0 v.x + M1 v.y) + M2 v.z), both additions fused."""
    out = np.empty(v.shape, np.float32)
    for r in range(3):
        out[:, r] = fma32(M[:, 2, r], v[:, 2], fma32(M[:, 1, r], v[:, 1], M[:, 0, r] * v[:, 0]))
    return out


def _tau_C(rec: np.ndarray):
    """tau and C of the records as WGSL matrices [n, col, row] (mpm.js:325-326)."""
    n = len(rec)
    tau = np.empty((n, 3, 3), np.float32)
    # mat3x3f(p.c0.w, p.f0.w, p.f1.w,  p.f0.w, p.c1.w, p.f2.w,  p.f1.w, p.f2.w, p.c2.w)
    cols = [(11, 23, 27), (23, 15, 31), (27, 31, 19)]
    for c, idx in enumerate(cols):
        for r, k in enumerate(idx):
            tau[:, c, r] = rec[:, k]
    C = np.stack([rec[:, 8:11], rec[:, 12:15], rec[:, 16:19]], axis=1).astype(np.float32)
    return tau, C


def p2g(rec: np.ndarray, mats: np.ndarray, vol4: float, dims, stats: dict | None = None) -> np.ndarray:
    """The P2G grid words (mpm.js:318-349): for each particle and each of its 27 stencil nodes,
    mom = wt * (mv + affine * dpos) with dpos = vec3f(off) - fx, wt = w[a].x * w[b].y * w[c].z,
    q = (toFixed(mom), toFixed(wt * mass)), added to the node with wrapping i32 arithmetic.
    Returns (X*Y*Z, 4) int32 in the reference layout (node = x + X*(y + Y*z)). stats (optional)
    receives "clamped": the number of contributions toFixed's clamp acted on."""
    rec = np.ascontiguousarray(rec, np.float32)
    X, Y, Z = (int(d) for d in dims)
    xp = rec[:, 0:3]
    base = np.floor(xp - F32(0.5)).astype(np.int64)
    fx = xp - base.astype(np.float32)
    w = bspline(fx)
    mid = rec[:, 3].view(np.uint32).astype(np.int64)
    mass = np.asarray(mats, np.float32).reshape(-1)[mid * 12 + 1]
    tau, C = _tau_C(rec)
    affine = F32(-np.float32(vol4)) * tau + mass[:, None, None] * C  # -vol4 * tau + mass * C
    mv = mass[:, None] * rec[:, 4:7]
    acc = np.zeros((X * Y * Z, 4), np.int64)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                off = np.array([a, b, c], np.float32)
                dpos = off[None, :] - fx
                wt = (w[a][:, 0] * w[b][:, 1]) * w[c][:, 2]
                mom = wt[:, None] * (mv + mat_vec(affine, dpos))
                if stats is not None:
                    stats["clamped"] = stats.get("clamped", 0) + int((np.abs(mom * FP) > F32(2.0e9)).sum())
                q = np.concatenate([to_fixed(mom), to_fixed(wt * mass)[:, None]], axis=1)
                g = base + np.array([a, b, c])
                node = g[:, 0] + X * (g[:, 1] + Y * g[:, 2])
                np.add.at(acc, node, q)
    return (((acc + (1 << 31)) % (1 << 32)) - (1 << 31)).astype(np.int32)


def g2p_velocity(rec: np.ndarray, gv_at, fused: bool = True) -> tuple[np.ndarray, np.ndarray]:
    """G2P's gathered v and C (mpm.js:483-497), node by node in the loop order a, b, c:
    v += wt * gv and B += outer3(wt * gv, dpos), C = 4 B. fused: the two readings a WGSL compiler
    may take of the accumulations, v = fma(wt, gv, v) and B_cr = fma(wgv_r, dpos_c, B_cr) with
    wgv = wt * gv, or every product rounded before its addition. gv_at(gx, gy, gz) -> (n, 3) float32
    node velocities. Returns v (n, 3), C [n, col, row]."""
    xp = np.ascontiguousarray(rec[:, 0:3], np.float32)
    base = np.floor(xp - F32(0.5)).astype(np.int64)
    fx = xp - base.astype(np.float32)
    w = bspline(fx)
    n = len(rec)
    v = np.zeros((n, 3), np.float32)
    B = np.zeros((n, 3, 3), np.float32)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                off = np.array([a, b, c], np.float32)
                dpos = off[None, :] - fx
                wt = (w[a][:, 0] * w[b][:, 1]) * w[c][:, 2]
                g = base + np.array([a, b, c])
                gv = gv_at(g[:, 0], g[:, 1], g[:, 2])
                wgv = wt[:, None] * gv
                v = fma32(wt[:, None], gv, v) if fused else v + wgv
                for col in range(3):
                    for row in range(3):
                        if fused:
                            B[:, col, row] = fma32(wgv[:, row], dpos[:, col], B[:, col, row])
                        else:
                            B[:, col, row] = B[:, col, row] + wgv[:, row] * dpos[:, col]
    return v, F32(4.0) * B
