"""Consumers of the particle layout and of the sort: readPositions / the render layers' position and
material slots (mpm.js:624-679) computed on the device, the device stream pointers, and SpatialGrid
build + 27-cell neighbour ranges (src/prims/grid.js:24-117) against a numpy restatement."""
import ctypes

import numpy as np
import pytest

from helpers import bits
from shaderintel_amd.mpm import MPM, REC_SLOT, spatial_grid

pytestmark = pytest.mark.gpu


def test_read_points_matches_records(ctx):
    print("[gpu] read_points vs read_particles", flush=True)
    sim = MPM(ctx, grid=48, size=0.75, origin=[0.1, -0.2, 0.3], maxParticles=50000, substeps=8)
    sim.add({"box": [[0.2, 0.1, 0.4], [0.5, 0.4, 0.7]]}, material="water", velocity=[1.0, 0.0, -0.5])
    sim.add({"sphere": [[0.6, 0.4, 0.6], 0.1]}, material="snow")
    sim.step()
    raw = sim.read_particles()
    pts, vel = sim.read_points(velocity=True)
    want = (raw[:, :3].astype(np.float64) * sim.dx + np.asarray(sim.origin)).astype(np.float32)
    assert np.array_equal(bits(pts[:, :3]), bits(want))
    assert np.array_equal(pts[:, 3].view(np.uint32), raw[:, 3].view(np.uint32))
    vs = sim.dx / sim.substepDt
    assert np.array_equal(bits(vel[:, :3]), bits((raw[:, 4:7].astype(np.float64) * vs).astype(np.float32)))
    assert np.array_equal(bits(vel[:, 3]), bits(raw[:, 7]))
    assert np.array_equal(bits(sim.readPositions()), bits(want))
    sim.destroy()


def test_particle_streams_hold_the_records(ctx):
    print("[gpu] device stream pointers", flush=True)
    sim = MPM(ctx, grid=32, maxParticles=20000, substeps=8)
    sim.add({"box": [[0.2, 0.2, 0.2], [0.5, 0.5, 0.5]]}, material="jelly")
    sim.step()
    raw = sim.read_particles()
    ptrs, cap = sim.particle_streams()
    assert cap >= sim.count and all(ptrs)
    hip = ctypes.CDLL("libamdhip64.so")
    n = sim.count
    streams = np.zeros((8, n, 4), np.float32)
    for k in range(8):
        assert hip.hipMemcpy(streams[k].ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptrs[k]),
                             ctypes.c_size_t(n * 16), 2) == 0  # hipMemcpyDeviceToHost
    assert sorted(REC_SLOT) == list(range(32))
    for f, slot in enumerate(REC_SLOT):
        assert np.array_equal(bits(streams[slot // 4, :, slot % 4]), bits(raw[:, f])), f"record float {f}"
    # positions and the material slot where the render layers read them (mpm.js:642-679)
    assert np.array_equal(bits(streams[0]), bits(raw[:, 0:4]))
    sim.destroy()


def _grid_reference(points, lo, hi, cell, queries):
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    dims = np.array([max(1, int(np.ceil((hi[a] - lo[a]) / cell))) for a in range(3)])
    inv = np.float32(1.0 / cell)

    def coord(p):
        c = np.floor((p.astype(np.float32) - lo.astype(np.float32)) * inv).astype(np.int64)
        return np.clip(c, 0, dims - 1)
    c = coord(points)
    cells = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
    order = np.argsort(cells, kind="stable").astype(np.uint32)
    start = np.concatenate([[0], np.cumsum(np.bincount(cells, minlength=int(np.prod(dims))))]).astype(np.uint32)
    nbrs, offs = [], [0]
    for q in coord(queries):
        x0, x1 = max(q[0] - 1, 0), min(q[0] + 1, dims[0] - 1)
        for z in range(max(q[2] - 1, 0), min(q[2] + 1, dims[2] - 1) + 1):
            for y in range(max(q[1] - 1, 0), min(q[1] + 1, dims[1] - 1) + 1):
                row = dims[0] * (y + dims[1] * z)
                nbrs.extend(range(start[row + x0], start[row + x1 + 1]))
        offs.append(len(nbrs))
    return order, start, np.asarray(offs, np.uint32), np.asarray(nbrs, np.uint32)


@pytest.mark.parametrize("n,m,cell", [(0, 5, 0.1), (1, 1, 0.5), (5000, 300, 0.07), (20000, 1000, 0.031)])
def test_spatial_grid_matches_reference(ctx, n, m, cell):
    print(f"[gpu] spatial grid n={n} m={m}", flush=True)
    rng = np.random.default_rng(n + m)
    pts = rng.uniform(-0.1, 1.1, (n, 3)).astype(np.float32)  # some outside: clamped to edge cells
    qs = rng.uniform(0.0, 1.0, (m, 3)).astype(np.float32)
    got = spatial_grid(ctx, pts, [0, 0, 0], [1, 1, 1], cell, qs)
    want = _grid_reference(pts, [0, 0, 0], [1, 1, 1], cell, qs)
    for g, w, name in zip(got, want, ("order", "start", "offsets", "neighbors")):
        assert np.array_equal(g, w), name
