"""The reference's own MPM tests (tests/physics.test.js:6-60) run through the HIP path, written as the
reference writes them (gpu.mpm, add, P.read / P.write, p2g, grid.read, step, _sort), with the
reference's assertions and tolerances, plus bit-exact checks against the oracle and the node-by-node
numpy P2G where the reference has none."""
from __future__ import annotations

import numpy as np
import pytest

import literal
import oracle_mpm as O
from helpers import bits, mulberry32

pytestmark = pytest.mark.gpu


def _oracle_of(sim, raw_records: np.ndarray) -> O.Scene:
    sc = O.Scene(sim.dims, len(raw_records), sim.ppcAxis, sim.sortEvery)
    sc.set_materials(sim.material_table())
    sc.write(raw_records)
    sc.set_count(len(raw_records))
    sc.step_count = sim.stepCount
    return sc


def test_p2g_conserves_mass_and_momentum(ctx):
    """physics.test.js:6-32: randomised v, C and tau (mulberry32 seed 3, in the reference's order);
    grid mass within 1e-4 relative, momentum within 1e-3 |p| + n 27 4 / 2^22."""
    print("[gpu] reference physics: P2G conservation (grid 48)", flush=True)
    sim = ctx.mpm({"grid": 48, "maxParticles": 60000, "substeps": 4})
    sim.add({"box": [[0.2, 0.2, 0.2], [0.7, 0.6, 0.7]]}, {"material": "water"})
    n = sim.count
    raw = sim.P.read(n * 8)
    R = mulberry32(3)
    for k in range(n):
        o = 32 * k
        for a in range(3):
            raw[o + 4 + a] = (R() - 0.5) * 0.4
        for c in range(3):
            for a in range(3):
                raw[o + 8 + 4 * c + a] = (R() - 0.5) * 0.2
        raw[o + 11] = (R() - 0.5) * 0.3; raw[o + 15] = (R() - 0.5) * 0.3; raw[o + 19] = (R() - 0.5) * 0.3
        raw[o + 23] = (R() - 0.5) * 0.3; raw[o + 27] = (R() - 0.5) * 0.3; raw[o + 31] = (R() - 0.5) * 0.3
    sim.P.write(raw)
    sim.p2g()
    g = sim.grid.read().astype(np.int64)
    FP, mass = 2.0 ** 22, 1.0 / 8.0
    gm = g[3::4].sum() / FP
    gp = np.array([g[0::4].sum(), g[1::4].sum(), g[2::4].sum()]) / FP
    rec = raw.reshape(-1, 32)
    pm = mass * n
    pp = mass * rec[:, 4:7].astype(np.float64).sum(0)
    rel_m = abs(gm - pm) / pm
    err_p = np.abs(gp - pp).max()
    assert rel_m < 1e-4, "mass not conserved"
    assert err_p < 1e-3 * max(1.0, float(np.linalg.norm(pp))) + n * 27 * 4 / FP, "momentum not conserved"
    # beyond the reference: the words themselves, against the oracle and the node-by-node reading
    sc = _oracle_of(sim, rec)
    sc.p2g()
    words = g.reshape(-1, 4).astype(np.int32)
    assert np.array_equal(words, sc.grid())
    assert np.array_equal(words, literal.p2g(rec, sim.material_table(), 0.5, sim.dims))
    sim.destroy()


def test_tiled_p2g_bit_identical_to_global(ctx):
    """physics.test.js:33-42, unchanged: three frames, _sort, then P2G into a second grid array with
    the global-atomic path and into the simulation's grid with the tiled path: every word equal."""
    print("[gpu] reference physics: tiled == global (grid 48)", flush=True)
    sim = ctx.mpm({"grid": 48, "maxParticles": 80000, "substeps": 4})
    sim.add({"box": [[0.1, 0.1, 0.1], [0.9, 0.4, 0.9]]}, {"material": "water"})
    for _ in range(3):
        sim.step()
    sim._sort()
    other = ctx.array("vec4i", sim.numNodes)
    sim.p2g(sim.P, True, other)
    sim.p2g(sim.P, False, sim.grid)
    a, b = other.read(), sim.grid.read()
    bad = np.flatnonzero(a != b)
    assert len(bad) == 0, f"word {bad[0]} differs"
    assert (a != 0).sum() > 1000
    sim.destroy()


def test_dam_break_stable_bounded_volume_preserved(ctx):
    """physics.test.js:43-60: water box + jelly sphere on a 48^3 grid, 150 frames of 8 substeps: no
    NaN, every particle inside [1, dims - 1], mean water J within 0.05 of 1. The first 10 frames are
    also checked bit for bit against the oracle."""
    print("[gpu] reference physics: dam break, 150 frames (grid 48)", flush=True)
    sim = ctx.mpm({"grid": 48, "maxParticles": 200000, "substeps": 8})
    sim.add({"box": [[0.05, 0.05, 0.05], [0.4, 0.6, 0.95]]}, {"material": "water"})
    sim.add({"sphere": [[0.7, 0.4, 0.5], 0.12]}, {"material": "jelly"})
    sc = _oracle_of(sim, sim.read_particles())
    g = sim.lattice_gravity()
    for f in range(150):
        sim.step()
        if f < 10:
            sc.step(8, g)
        if f == 9:
            assert np.array_equal(bits(sim.read_particles()), bits(sc.read(sim.count))), "frame 10 differs from the oracle"
    n = sim.count
    raw = sim.P.read(n * 8).reshape(-1, 32)
    x = raw[:, 0:3]
    nan = int((~np.isfinite(x)).sum())
    out = int(((x < 1) | (x > np.asarray(sim.dims) - 1)).sum())
    water = raw[:, 3].view(np.uint32) == 0
    mean_j = float(raw[water, 7].astype(np.float64).mean())
    assert nan == 0 and out == 0, "unstable"
    assert abs(mean_j - 1) < 0.05, f"volume drift (mean J {mean_j})"
    sim.destroy()
