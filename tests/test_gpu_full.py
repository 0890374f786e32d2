"""The 8,388,608-particle four-material scene on its 256^3 grid (water, jelly, snow, Drucker-Prager
sand): the measured workload at its full size. Twelve substeps (three sorts, nine fused G2P + P2G
launches) bit for bit against the oracle, the grid update's momentum sums bit for bit and the
momentum law they state, then the conservation law of the P2G grid at that size."""
import numpy as np
import pytest

import oracle_mpm as O
from helpers import bits
from shaderintel_amd import scenes
from shaderintel_amd.mpm import MPM, load_library, _ptr

pytestmark = pytest.mark.gpu

SUBSTEPS = 12


@pytest.fixture(scope="module")
def full(ctx):
    print(f"[gpu] full size: mixed256, {scenes.expected_count('mixed256')} particles, {SUBSTEPS} substeps",
          flush=True)
    sim = MPM(ctx, **scenes.options("mixed256"))
    n = scenes.populate(sim, "mixed256")
    rec = sim.read_particles()
    sc = O.Scene(sim.dims, n, sim.ppcAxis, 4)
    sc.set_materials(sim.material_table())
    sc.write(rec)
    sc.set_count(n)
    del rec
    g = sim.lattice_gravity()
    sim.set_grid_law(True)
    sc.set_grid_law(True)
    assert load_library().mpm_step(sim._h, SUBSTEPS, _ptr(g), 0.4, 0.9) == 0
    sc.step(SUBSTEPS, g)
    yield sim, sc, n
    sim.destroy()


def test_full_scene_matches_oracle(full):
    sim, sc, n = full
    got = sim.read_particles()
    want = sc.read(n)
    bad = np.argwhere(bits(got) != bits(want))
    assert len(bad) == 0, f"particles: {len(bad)} words differ, first {tuple(bad[0])}"
    del got, want
    bad = np.argwhere(bits(sim.read_grid()) != bits(sc.grid()))
    assert len(bad) == 0, f"grid: {len(bad)} words differ, first {tuple(bad[0])}"


def test_full_scene_momentum_law(full):
    """The grid update's sums over the twelve substeps are the oracle's, word for word, and the grid
    momentum after the updates minus the P2G momentum is the gravity plus boundary impulse within
    1e-5 of the impulses."""
    sim, sc, n = full
    law = sim.read_grid_law()
    want = sc.read_grid_law()
    assert law["raw"] == [int(np.int64(x)) for x in want[:14].astype(np.int64)]
    assert law["updates"] == SUBSTEPS == int(want[14])
    G, B = law["gravity"], law["boundary"]
    resid = law["momentum"] - law["p2g"] - G - B
    assert np.abs(resid).max() <= 1e-5 * (np.abs(G).sum() + np.abs(B).sum()), (resid, G, B)
    assert B[1] > 0  # the floor holds the blocks up


def test_full_scene_conservation(full):
    """verify/index.js:322-329 (law mpm-conservation) with checkMPM's tolerance (:175-193), at full
    size: the P2G grid's total mass and momentum against the particles' (four materials, so each
    particle's mass comes from its material row)."""
    sim, sc, n = full
    raw = sim.read_particles()
    mass = np.abs(sim.material_table()[:, 1].astype(np.float64))[raw[:, 3].view(np.uint32)]
    pm = mass.sum()
    pv = (mass[:, None] * raw[:, 4:7].astype(np.float64)).sum(0)
    del raw
    sim.p2g()
    g = sim.read_grid().astype(np.float64)
    FP = 2.0 ** 22
    gm = g[:, 3].sum() / FP
    gp = g[:, :3].sum(0) / FP
    mass_err = abs(gm - pm) / pm
    mom_err = np.linalg.norm(gp - pv) / pm
    assert mass_err < 1e-4 and mom_err < 1e-4 + n * 27 * 4 / FP / pm, (mass_err, mom_err)
