"""The momentum law through the HIP path: the grid update's sums (mpm_read_grid_law) bit for bit
against the oracle's, and the law they state: per frame the grid momentum after the updates minus
the P2G momentum equals the gravity impulse plus the boundary impulse of the walls, the Coulomb floor
and a moving sphere collider within 1e-5 of the impulses, and the update's mass words are exactly
the P2G grid's (checkMPM's mass law, src/verify/index.js:175-193, carried through the update)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_mpm as O
from helpers import bits
from shaderintel_amd.mpm import MPM, load_library, _ptr

pytestmark = pytest.mark.gpu

DX = 1 / 64


def _sim(ctx):
    sim = MPM(ctx, grid=64, maxParticles=70000, substeps=8)
    sim.add({"box": [[3 * DX, 3 * DX, 3 * DX], [11 * DX, 35 * DX, 35 * DX]]}, material="water", velocity=[1.0, 0, 0],
            jitter=1.0)
    # a sphere of radius 5 cells moving toward the block at 0.05 cells per substep
    sim.setColliders([{"type": "sphere", "center": [20 * DX, 6 * DX, 19 * DX], "radius": 5 * DX, "friction": 0.3,
                       "velocity": [-0.05 * DX / sim.substepDt, 0, 0]}])
    rec = sim.read_particles()
    sc = O.Scene(sim.dims, len(rec), sim.ppcAxis, 4)
    sc.set_materials(sim.material_table())
    sc.set_colliders(sim.collider_table(), 1)
    sc.write(rec)
    sc.set_count(len(rec))
    return sim, sc, len(rec)


def _raw(oracle_law: np.ndarray) -> list:
    return [int(x) for x in oracle_law[:14].astype(np.int64)]


def test_grid_law_matches_oracle_and_holds(ctx):
    print("[gpu] grid law: dam break + moving sphere collider, 5 frames of 8 substeps", flush=True)
    sim, sc, n = _sim(ctx)
    g = sim.lattice_gravity()
    sim.set_grid_law(True)
    sc.set_grid_law(True)
    bnd = np.zeros(3)
    for frame in range(5):
        sim.step()
        sc.step(8, g)
        law = sim.read_grid_law(reset=True)
        want = sc.read_grid_law(reset=True)
        assert law["raw"] == _raw(want), frame
        assert law["updates"] == 8 == int(want[14])
        G, B = law["gravity"], law["boundary"]
        resid = law["momentum"] - law["p2g"] - G - B
        assert np.abs(resid).max() <= 1e-5 * (np.abs(G).sum() + np.abs(B).sum()), (frame, resid)
        bnd += B
    assert bnd[0] < 0 and bnd[1] > 0  # the collider pushes back along -x, the floor up
    got, exp = sim.read_particles(), sc.read(n)
    assert np.array_equal(bits(got), bits(exp))
    sim.destroy()


def test_grid_law_mass_words_are_the_p2g_grid(ctx):
    print("[gpu] grid law: mass and momentum words against the P2G grid", flush=True)
    sim, _, _ = _sim(ctx)
    g = sim.lattice_gravity()
    L = load_library()
    for _ in range(3):
        assert L.mpm_step(sim._h, 3, _ptr(g), 0.4, 0.9) == 0
        sim.p2g()
        words = sim.read_grid().astype(np.int64)
        sim.set_grid_law(True)
        assert L.mpm_step(sim._h, 1, _ptr(g), 0.4, 0.9) == 0
        law = sim.read_grid_law()
        sim.set_grid_law(False)
        assert law["raw"][0] == int(words[:, 3].sum())
        assert law["raw"][1:4] == [int(x) for x in words[:, :3].sum(0)]
        assert law["updates"] == 1
    sim.destroy()


def test_grid_law_off_by_default_and_result_neutral(ctx):
    """The sums are off until enabled, and enabling them changes no particle or grid bit."""
    print("[gpu] grid law: neutral", flush=True)
    a, _, _ = _sim(ctx)
    b, _, _ = _sim(ctx)
    assert a.read_grid_law()["updates"] == 0
    b.set_grid_law(True)
    for _ in range(2):
        a.step()
        b.step()
    assert a.read_grid_law()["updates"] == 0 and b.read_grid_law()["updates"] == 16
    assert np.array_equal(bits(a.read_particles()), bits(b.read_particles()))
    assert np.array_equal(a.read_grid(), b.read_grid())
    a.destroy()
    b.destroy()
