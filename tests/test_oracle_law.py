"""The momentum law of a substep on the oracle (CPU): the grid update is the only place momentum
enters or leaves, so for every substep the grid momentum after the update minus the P2G momentum
equals the gravity impulse sum(m g) plus the boundary impulse of the walls, the Coulomb floor and the
colliders (mpm.js:428-470), within 1e-5 of the impulses; the mass words of that update are those of
the P2G grid (checkMPM's mass law, src/verify/index.js:175-193); and over the whole substep the
particles' momentum changes by the same impulses, within 1e-5 of their momentum. The boundary
impulse is pinned on hand-built nodes by the closed forms of the floor and wall rules."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_mpm as O

FP = float(1 << 22)
DX, DT = 1 / 64, 1 / 480


def _law(raw: np.ndarray) -> dict:
    L = raw.astype(np.int64)
    return {"mass_words": int(L[0]), "p2g": L[1:4] / FP, "momentum": L[4:7] / 2.0 ** 32,
            "gravity": L[7:10] / 2.0 ** 32, "boundary": L[10:13] / 2.0 ** 32, "nodes": int(L[13]),
            "updates": int(raw[14])}


def _dam_with_collider(n_cells=(8, 32, 32)):
    dims = [64, 64, 64]
    hi = [3 + n_cells[0], 3 + n_cells[1], 3 + n_cells[2]]
    pts = O.sample_shape(0, [3 * DX, 3 * DX, 3 * DX] + [h * DX for h in hi], DX / 2, 1.0, 1)
    rec = O.pack_particles(pts, [0, 0, 0], DX, dims, 3.0, 0, [1.0, 0, 0], DT / DX)  # 1 m/s toward +x
    tbl = np.zeros((8, 12), np.float32)
    tbl[0] = O.material_row({"kind": 0, "density": 1.0, "stiffness": 1.0, "viscosity": 0.0}, 2)
    col = np.zeros(16 * 12, np.float32)
    col.view(np.uint32)[0] = 0  # a sphere of radius 5 cells moving toward -x, friction 0.3
    col[1:4] = [20, 6, 19]
    col[4] = 5.0
    col[7] = 0.3
    col[8:11] = [-0.05, 0, 0]
    return dims, rec, tbl, col


def _scene(dims, rec, tbl, col):
    sc = O.Scene(dims, len(rec), 2, 4)
    sc.set_materials(tbl)
    sc.write(rec)
    sc.set_count(len(rec))
    sc.set_colliders(col, 1)
    return sc


def test_substep_momentum_law_dam_break_with_collider():
    dims, rec, tbl, col = _dam_with_collider()
    sc = _scene(dims, rec, tbl, col)
    probe = _scene(dims, rec, tbl, col)  # the same state, for reading the P2G grid of each substep
    g = np.array([0.0, np.float32(-9.81 * DT * DT / DX), 0.0], np.float32)
    n, m = len(rec), float(tbl[0, 1])
    sc.set_grid_law(True)
    wall = floor = 0.0
    for k in range(40):
        cur = sc.read(n)
        probe.write(cur)
        probe.step_count = sc.step_count
        if sc.step_count % 4 == 0:
            probe.sort()  # the substep sorts before its P2G; P2G sums do not depend on the order
        probe.p2g()
        grid = probe.grid().astype(np.int64)
        before = m * cur[:, 4:7].astype(np.float64).sum(0)
        sc.step(1, g)
        after_rec = sc.read(n)
        after = m * after_rec[:, 4:7].astype(np.float64).sum(0)
        assert np.abs(after_rec[:, 4:7]).max() < 0.9  # clear of the speed clamp (mpm.js:501)
        L = _law(sc.read_grid_law(reset=True))
        assert L["updates"] == 1
        # mass: the update's mass words are the P2G grid's, and within checkMPM's fixed-point bound of
        # the particles' mass
        assert L["mass_words"] == int(grid[:, 3].sum())
        assert abs(L["mass_words"] / FP - n * m) <= n * 27 / FP
        assert np.array_equal(np.rint(L["p2g"] * FP).astype(np.int64), grid[:, :3].sum(0))
        G, B = L["gravity"], L["boundary"]
        assert np.allclose(G, np.array([0, g[1], 0]) * L["mass_words"] / FP, rtol=1e-6, atol=1e-9)
        scale = np.abs(G).sum() + np.abs(B).sum()
        assert np.abs(L["momentum"] - L["p2g"] - G - B).max() <= 1e-5 * scale, k
        # the particles: their momentum changes by the same impulses, within 1e-5 of their momentum
        # (the P2G transfer's fixed-point rounding is the largest part of the difference)
        pscale = m * np.abs(cur[:, 4:7].astype(np.float64)).sum()
        assert np.abs((after - before) - G - B).max() <= 1e-5 * pscale, k
        wall, floor = wall + B[0], floor + B[1]
    assert wall < -1.0 and floor > 1.0  # the collider pushes the block back (-x), the floor up (+y)


def test_law_accumulates_over_substeps_and_sorts():
    """The sums over 12 substeps in one call equal the sums of 12 single substeps."""
    dims, rec, tbl, col = _dam_with_collider((8, 16, 16))
    g = np.array([0.0, np.float32(-9.81 * DT * DT / DX), 0.0], np.float32)
    a, b = _scene(dims, rec, tbl, col), _scene(dims, rec, tbl, col)
    a.set_grid_law(True)
    b.set_grid_law(True)
    a.step(12, g)
    tot = np.zeros(16, np.uint64)
    for _ in range(12):
        b.step(1, g)
        tot += b.read_grid_law(reset=True)
    assert np.array_equal(a.read_grid_law(), tot)
    assert int(tot[14]) == 12


def _grid_with(node_vals, dims):
    w = np.zeros((dims[0] * dims[1] * dims[2], 4), np.int32)
    for (x, y, z), (p, m) in node_vals.items():
        i = x + dims[0] * (y + dims[1] * z)
        w[i, :3] = np.rint(np.asarray(p) * FP)
        w[i, 3] = np.rint(m * FP)
    return w


def test_boundary_impulse_closed_forms():
    """Hand-built nodes: an interior node takes no boundary impulse; a node under the floor
    (y < bound) moving down loses its downward momentum and, by Coulomb friction, a share
    0.4 |vn| / |vt| of its tangential momentum; a node beyond the +x wall moving out loses its x
    momentum."""
    dims = [16, 16, 16]
    g = np.array([0.0, -0.0027, 0.0], np.float32)
    m = 0.5
    vals = {(8, 8, 8): ((0.05, 0.02, -0.01), m),   # interior
            (8, 1, 8): ((0.04, -0.03, 0.0), m),    # floor, moving down
            (14, 8, 8): ((0.06, 0.0, 0.01), m)}    # past the +x wall (x > 16 - 3), moving out
    sc = O.Scene(dims, 1, 2, 4)
    sc.set_grid_law(True)
    sc.write_grid(_grid_with(vals, dims))
    sc.grid_update(g, 0.4)
    L = _law(sc.read_grid_law())
    vf = {k: np.asarray(p) / mm + g.astype(np.float64) for k, (p, mm) in vals.items()}
    # floor: vn = -vy, v.y -> 0, tangential scale f = max(0, 1 - 0.4 vn / vt)
    v = vf[(8, 1, 8)]
    vt = np.hypot(v[0], v[2])
    f = max(0.0, 1 - 0.4 * (-v[1]) / vt)
    floor_imp = m * (np.array([v[0] * f, 0.0, v[2] * f]) - v)
    wall_imp = m * np.array([-vf[(14, 8, 8)][0], 0.0, 0.0])
    np.testing.assert_allclose(L["boundary"], floor_imp + wall_imp, rtol=1e-5, atol=1e-8)
    assert floor_imp[1] > 0 and floor_imp[0] < 0 and wall_imp[0] < 0
    assert L["nodes"] == 3
    np.testing.assert_allclose(L["gravity"], 3 * m * g.astype(np.float64), rtol=1e-6)
    np.testing.assert_allclose(L["momentum"], L["p2g"] + L["gravity"] + L["boundary"], atol=1e-6)
