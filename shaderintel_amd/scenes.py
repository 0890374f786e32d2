"""Synthetic MLS-MPM workloads, stated through the reference's world API.

Every scene uses the reference's default regime: ppcAxis 2 (8 particles per cell), bound 3,
sortEvery 4, 8 substeps per 1/60 s frame (dt = 1/480 s), no colliders, and a world size of
grid/64 m so that dx = 1/64 m and the lattice gravity is -9.81*dt^2/dx for every scene. Boxes are
given in cells as [lo, hi). Seeds follow MPM.add's default (count + 1, mpm.js:210) with jitter 1.

- ``dam64``: 64^3 grid, water box x in [3,11), y in [3,35), z in [3,35): 65,536 particles.
- ``mixed128``: 128^3 grid, water [3,67)x[3,35)x[3,35) and a jelly box [72,104)x[40,72)x[40,104)
  falling at 2 cells/s: 1,048,576 particles.
- ``mixed256``: 256^3 grid, four 64^3-cell blocks at y in [3,67): water, jelly, snow and
  Drucker-Prager sand (friction 30 deg), Gaussian velocity noise 0.1 cells/s: 8,388,608 particles.
- ``mixed256-water`` / ``-jelly`` / ``-snow`` / ``-sand``: the same four blocks, all of one material.
"""
from __future__ import annotations

import numpy as np

from . import mpm as _m

SAND30 = {"kind": "sand", "density": 1.6, "stiffness": 0.6, "poisson": 0.3, "friction": 30}

SCENES = {
    "dam64": {
        "grid": 64,
        "blocks": [("water", (3, 3, 3), (11, 35, 35), (0, 0, 0), 0.0)],
    },
    "mixed128": {
        "grid": 128,
        "blocks": [("water", (3, 3, 3), (67, 35, 35), (0, 0, 0), 0.0),
                   ("jelly", (72, 40, 40), (104, 72, 104), (0, -2, 0), 0.0)],
    },
    "mixed256": {
        "grid": 256,
        "blocks": [("water", (3, 3, 3), (67, 67, 67), (0, 0, 0), 0.1),
                   ("jelly", (128, 3, 3), (192, 67, 67), (0, 0, 0), 0.1),
                   ("snow", (3, 3, 128), (67, 67, 192), (0, 0, 0), 0.1),
                   (SAND30, (128, 3, 128), (192, 67, 192), (0, 0, 0), 0.1)],
    },
}

# The 8M scene with all four blocks of one material (same boxes, noise and seeds): the per-material
# cost of the substep (the snow and sand SVD against water's and jelly's closed forms).
for _mat, _key in (("water", "water"), ("jelly", "jelly"), ("snow", "snow"), (SAND30, "sand")):
    SCENES["mixed256-" + _key] = {"grid": 256, "blocks": [(_mat,) + b[1:] for b in SCENES["mixed256"]["blocks"]]}

SUBSTEPS = 8
DX = 1.0 / 64.0


def options(name: str) -> dict:
    sc = SCENES[name]
    n = expected_count(name)
    return {"grid": sc["grid"], "size": sc["grid"] / 64.0, "maxParticles": n, "substeps": SUBSTEPS, "sortEvery": 4,
            "ppcAxis": 2}


def expected_count(name: str) -> int:
    return sum(8 * int(np.prod(np.subtract(hi, lo))) for _, lo, hi, _, _ in SCENES[name]["blocks"])


def populate(sim: "_m.MPM", name: str) -> int:
    """Add the scene's blocks to a simulation through MPM.add; returns the particle count."""
    for mat, lo, hi, vel, noise in SCENES[name]["blocks"]:
        box = [[l * DX for l in lo], [h * DX for h in hi]]
        sim.add({"box": box}, material=mat, velocity=[v * DX for v in vel], jitter=1.0, noise=noise * DX)
    return sim.count


def host_blocks(grid: int, blocks, ppc: int = 2, gravity=(0.0, -9.81, 0.0), substeps: int = SUBSTEPS, dims=None):
    """A scene of material boxes built on the host only (no device), as MPM.add would build it on a
    grid^3 lattice (or `dims`) with dx = 1/64 m. blocks: (material, lo cells, hi cells, velocity cells/s, noise
    cells/s). Returns (dims, material table (8, 12), records (N, 32) float32, lattice gravity (3,)
    float32)."""
    import json
    dims = list(dims) if dims is not None else [grid] * 3
    dx = (grid / 64.0) / grid
    dt = (1 / 60) * 1 / substeps
    materials, index, recs = [], {}, []
    count = 0
    L = _m.load_library()
    for mat, lo, hi, vel, noise in blocks:
        d = {**_m.MATERIALS[mat], "name": mat} if isinstance(mat, str) else dict(mat)
        if isinstance(d["kind"], str):
            d["kind"] = _m.MPM_KINDS[d["kind"]]
        key = json.dumps(d)
        if key not in index:
            index[key] = len(materials)
            materials.append(d)
        mid = index[key]
        box = [[l * DX for l in lo], [h * DX for h in hi]]
        pts = _m.sample_shape({"box": box}, dx / ppc, 1.0, count + 1)
        rec = np.zeros((len(pts), 32), np.float32)
        L.mpm_pack_particles(_m._ptr(np.ascontiguousarray(pts)), len(pts), _m._ptr(np.zeros(3)), dx,
                             _m._ptr(np.asarray(dims, np.int32)), 3.0, mid,
                             _m._ptr(np.asarray([v * DX for v in vel], np.float64)), dt / dx, float(noise * DX),
                             ((count + 1) * 2654435761) & 0xFFFFFFFF, _m._ptr(rec))
        recs.append(rec)
        count += len(pts)
    tbl = np.zeros((8, 12), np.float32)
    for k, m in enumerate(materials):
        tbl[k] = _m.material_row(m, ppc)
    g = np.array([(x * dt * dt) / dx for x in gravity], np.float32)
    recs = np.concatenate(recs) if recs else np.zeros((0, 32), np.float32)
    return dims, tbl, recs, g


def host_build(name: str):
    """The named scene built on the host only (see host_blocks)."""
    sc = SCENES[name]
    return host_blocks(sc["grid"], sc["blocks"])


MATERIAL_CHOICES = ["water", "jelly", "snow", SAND30, "honey", "rubber"]


def random_blocks(rng: np.random.Generator, grid: int, max_blocks: int = 3, max_edge: int = 24):
    """1..max_blocks random material boxes inside a grid^3 lattice (cells, clear of the 3-cell
    boundary), with random velocities and noise: the per-request scenes of the batched workload."""
    blocks = []
    for _ in range(int(rng.integers(1, max_blocks + 1))):
        edge = rng.integers(4, min(max_edge, grid - 8) + 1, size=3)
        lo = [int(rng.integers(3, grid - 3 - e + 1)) for e in edge]
        hi = [l + int(e) for l, e in zip(lo, edge)]
        mat = MATERIAL_CHOICES[int(rng.integers(len(MATERIAL_CHOICES)))]
        vel = [float(v) for v in rng.uniform(-8, 8, size=3)]
        blocks.append((mat, lo, hi, vel, float(rng.uniform(0, 0.2))))
    return blocks


BATCH_EDGES = (32, 48, 64, 96, 128)
BATCH_MATERIALS = ("water", "jelly", "snow", SAND30)


def batched_blocks(seed: int = 1, count: int = 64):
    """The batched workload: `count` independent scenes drawn from one master seed. Each has a cubic
    grid with edge in {32, 48, 64, 96, 128}, one box of water, jelly, snow or sand (30 deg) holding N
    particles with N log-uniform in [1e4, 1e6] and at most 8 per cell of the interior (edge - 6)^3,
    placed at a random position with a random velocity (|v_i| <= 4 cells/s). Returns a list of
    (edge, blocks) for host_blocks."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        edge = int(BATCH_EDGES[int(rng.integers(len(BATCH_EDGES)))])
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        n = min(n, 8 * (edge - 6) ** 3)
        cells = max(1, n // 8)
        side = min(edge - 6, max(1, int(round(cells ** (1 / 3)))))
        height = min(edge - 6, max(1, int(round(cells / (side * side)))))
        size = (side, height, side)
        lo = [int(rng.integers(3, edge - 3 - s + 1)) for s in size]
        hi = [l + s for l, s in zip(lo, size)]
        mat = BATCH_MATERIALS[int(rng.integers(len(BATCH_MATERIALS)))]
        vel = tuple(float(v) for v in rng.uniform(-4, 4, 3))
        out.append((edge, [(mat, lo, hi, vel, 0.0)]))
    return out
