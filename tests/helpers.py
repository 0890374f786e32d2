"""Shared helpers for the test suite (seeded generators, comparisons)."""
from __future__ import annotations

import numpy as np


def mulberry32(seed: int = 1):
    """The reference test harness's PRNG (tests/harness.js:59-66), as a Python generator of floats."""
    state = [seed & 0xFFFFFFFF]

    def rnd() -> float:
        state[0] = (state[0] + 0x6D2B79F5) & 0xFFFFFFFF
        s = state[0]
        t = ((s ^ (s >> 15)) * (1 | s)) & 0xFFFFFFFF
        t = ((t + (((t ^ (t >> 7)) * (61 | t)) & 0xFFFFFFFF)) & 0xFFFFFFFF) ^ t
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0

    return rnd


def randomize_transfer_terms(raw: np.ndarray, seed: int = 3) -> np.ndarray:
    """tests/physics.test.js:9-17: random v, C and tau in every record (in place, returns raw)."""
    R = mulberry32(seed)
    n = raw.shape[0]
    for k in range(n):
        for a in range(3):
            raw[k, 4 + a] = (R() - 0.5) * 0.4
        for c in range(3):
            for a in range(3):
                raw[k, 8 + 4 * c + a] = (R() - 0.5) * 0.2
        raw[k, 11] = (R() - 0.5) * 0.3
        raw[k, 15] = (R() - 0.5) * 0.3
        raw[k, 19] = (R() - 0.5) * 0.3
        raw[k, 23] = (R() - 0.5) * 0.3
        raw[k, 27] = (R() - 0.5) * 0.3
        raw[k, 31] = (R() - 0.5) * 0.3
    return raw


def randomize_fast(raw: np.ndarray, seed: int = 3) -> np.ndarray:
    """Vectorised variant of randomize_transfer_terms for large arrays (numpy generator)."""
    rng = np.random.default_rng(seed)
    n = raw.shape[0]
    raw[:, 4:7] = (rng.random((n, 3)) - 0.5) * 0.4
    for c in range(3):
        raw[:, 8 + 4 * c:11 + 4 * c] = (rng.random((n, 3)) - 0.5) * 0.2
    for j in (11, 15, 19, 23, 27, 31):
        raw[:, j] = (rng.random(n) - 0.5) * 0.3
    return raw


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint32)
