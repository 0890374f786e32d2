"""Edge cases of the substep against the oracle, bit for bit: all three collider kinds (moving,
with friction), retiming (rescaleK), grids whose edges are not multiples of 4 and not cubic,
1 and 3 particles per cell axis, a sort every substep, a full particle buffer, an empty
simulation, and particles so sparse that the work items' tiles take arena space by atomics."""
import numpy as np
import pytest

import oracle_mpm as O
from helpers import bits
from shaderintel_amd.mpm import MPM, diagnostics, load_library, _ptr

pytestmark = pytest.mark.gpu


def _oracle_of(sim, rec, sort_every=4):
    sc = O.Scene(sim.dims, max(len(rec), 1), sim.ppcAxis, sort_every)
    sc.set_materials(sim.material_table())
    sc.write(rec)
    sc.set_count(len(rec))
    return sc


def _same(a, b, what):
    bad = np.argwhere(bits(a) != bits(b))
    assert len(bad) == 0, f"{what}: {len(bad)} words differ, first {tuple(bad[0])}"


def test_colliders_all_kinds(ctx):
    print("[gpu] colliders: sphere, box, plane (moving, friction)", flush=True)
    sim = MPM(ctx, grid=40, maxParticles=80000, substeps=8)
    sim.add({"box": [[0.15, 0.3, 0.15], [0.85, 0.7, 0.85]]}, material="water", velocity=[0.3, -1.5, 0.2], jitter=1.0)
    sim.add({"box": [[0.3, 0.72, 0.3], [0.6, 0.85, 0.6]]}, material="snow", velocity=[0, -2.0, 0], jitter=1.0)
    sim.setColliders([
        {"type": "sphere", "center": [0.5, 0.25, 0.5], "radius": 0.12, "friction": 0.4, "velocity": [0.2, 0.0, 0.0]},
        {"type": "box", "center": [0.25, 0.2, 0.7], "half": [0.08, 0.05, 0.1], "friction": 0.1},
        {"type": "plane", "center": [0.0, 0.12, 0.0], "normal": [0.0, 1.0, 0.0], "friction": 0.8,
         "velocity": [0.0, 0.05, 0.0]},
    ])
    rec = sim.read_particles()
    sc = _oracle_of(sim, rec)
    sc.set_colliders(sim.collider_table(), 3)
    sim.step()
    sim.step()
    sc.step(16, sim.lattice_gravity())
    _same(sim.read_particles(), sc.read(len(rec)), "particles")
    _same(sim.read_grid(), sc.grid(), "grid")
    sim.destroy()


def test_retime_rescale(ctx):
    print("[gpu] retime: substeps 8 -> 16 -> 5, timeScale 0.5", flush=True)
    sim = MPM(ctx, grid=32, maxParticles=40000, substeps=8)
    sim.add({"box": [[0.2, 0.2, 0.2], [0.6, 0.6, 0.6]]}, material="jelly", velocity=[1.0, 0.5, -0.5], jitter=1.0)
    rec = sim.read_particles()
    sc = _oracle_of(sim, rec)
    sim.step()
    sc.step(8, sim.lattice_gravity())
    for change in (lambda: setattr(sim, "substeps", 16), lambda: setattr(sim, "timeScale", 0.5),
                   lambda: setattr(sim, "substeps", 5)):
        before = sim.substepDt
        change()
        sc.rescale(np.float32(sim.substepDt / before))
        sim.step()
        sc.step(sim.substeps, sim.lattice_gravity())
        _same(sim.read_particles(), sc.read(len(rec)), f"after retime to {sim.substeps} substeps")
    sim.destroy()


@pytest.mark.parametrize("size,grid,ppc", [([1.0, 0.7, 0.55], 37, 2), ([0.6, 1.0, 0.8], 30, 1),
                                           ([1.0, 1.0, 1.0], 21, 3)])
def test_odd_grids_and_ppc(ctx, size, grid, ppc):
    print(f"[gpu] grid {grid} size {size} ppcAxis {ppc}", flush=True)
    sim = MPM(ctx, grid=grid, size=size, maxParticles=100000, substeps=8, ppcAxis=ppc)
    sim.add({"box": [[0.1, 0.1, 0.1], [0.5, 0.45, 0.4]]}, material="sand", velocity=[0.6, 0.0, 0.3], jitter=1.0)
    sim.add({"sphere": [[0.45, 0.3, 0.3], 0.1]}, material="rubber", velocity=[-0.5, 0.2, 0.0])
    rec = sim.read_particles()
    sc = _oracle_of(sim, rec)
    sim.step()
    sim.step()
    sc.step(16, sim.lattice_gravity())
    _same(sim.read_particles(), sc.read(len(rec)), "particles")
    _same(sim.read_grid(), sc.grid(), "grid")
    sim.destroy()


def test_sparse_items_take_arena_space_by_atomics(ctx):
    """10,000 particles scattered over a 128^3 grid, most alone in their 4^3-cell block: about 8,600
    work items, whose window-tile slots (1,000 nodes each) do not fit the 7.1M-node arena, so the P2G
    kernels take each tile's space from the arena's counters instead of a slot per item. Particles
    and grid equal the oracle's after two frames (with the slots, every other test's path)."""
    print("[gpu] sparse items: tiles placed by arena atomics", flush=True)
    sim = MPM(ctx, grid=128, maxParticles=12000, substeps=8)
    sim.add({"box": [[0.4, 0.4, 0.4], [0.45, 0.45, 0.45]]}, material="snow", velocity=[0.2, -0.1, 0.3], noise=0.2)
    base = sim.read_particles()
    n = 10000
    rng = np.random.default_rng(21)
    rec = base[rng.integers(0, len(base), n)].copy()
    rec[:, 0:3] = rng.uniform(4.0, 123.0, (n, 3)).astype(np.float32)
    sim.write_particles(rec)
    assert load_library().mpm_set_count(sim._h, n) == 0
    sc = _oracle_of(sim, rec)
    sim.step()
    sim.step()
    sc.step(16, sim.lattice_gravity())
    d = diagnostics(sim._h)
    assert d["work_items"] * 1000 > d["tile_nodes_capacity"] >= d["tile_nodes_used"] > 0, d
    _same(sim.read_particles(), sc.read(n), "particles")
    _same(sim.read_grid(), sc.grid(), "grid")
    sim.destroy()


def test_window_changes_between_calls(ctx):
    """The P2G window (MPM_TUNE_DRIFT) switched between calls, mid sort period: S and the grid update's
    candidate tiles follow the window (a drift-3 tile reaches one node block further than a drift-2
    one), so a switch rebuilds the work items and S before the next P2G. Particles and grid equal the
    oracle's after every call."""
    print("[gpu] window changes between calls", flush=True)
    sim = MPM(ctx, grid=40, maxParticles=120000, substeps=8)
    sim.add({"box": [[0.1, 0.1, 0.1], [0.45, 0.45, 0.45]]}, material="water", velocity=[0.8, -0.5, 0.3], jitter=1.0)
    sim.add({"box": [[0.55, 0.2, 0.55], [0.9, 0.55, 0.9]]}, material="sand", velocity=[-0.2, -2.0, 0.1], jitter=1.0)
    rec = sim.read_particles()
    sc = _oracle_of(sim, rec)
    g = sim.lattice_gravity()
    L = load_library()
    for k, drift in ((6, 2), (5, 3), (7, 2), (3, 3)):
        sim.set_tuning(drift=drift)
        assert L.mpm_step(sim._h, k, _ptr(g), 0.4, 0.9) == 0
        sc.step(k, g)
        _same(sim.read_particles(), sc.read(len(rec)), f"particles after {k} substeps at drift {drift}")
        _same(sim.read_grid(), sc.grid(), f"grid after {k} substeps at drift {drift}")
    sim.destroy()


def test_sort_every_substep_and_full_buffer(ctx):
    print("[gpu] sortEvery 1, buffer exactly full", flush=True)
    probe = MPM(ctx, grid=32, maxParticles=100000, substeps=8)
    probe.add({"box": [[0.2, 0.2, 0.2], [0.5, 0.5, 0.5]]}, material="honey", velocity=[0.0, -1.0, 0.5])
    rec = probe.read_particles()
    tbl = probe.material_table()
    probe.destroy()
    sim = MPM(ctx, grid=32, maxParticles=len(rec), substeps=8, sortEvery=1)
    L = load_library()
    assert L.mpm_set_materials(sim._h, _ptr(tbl), 8) == 0
    sim.write_particles(rec)
    assert L.mpm_set_count(sim._h, len(rec)) == 0
    sc = O.Scene(sim.dims, len(rec), 2, 1)
    sc.set_materials(tbl)
    sc.write(rec)
    sc.set_count(len(rec))
    g = sim.lattice_gravity()
    assert L.mpm_step(sim._h, 12, _ptr(g), 0.4, 0.9) == 0
    sc.step(12, g)
    _same(sim.read_particles(), sc.read(len(rec)), "particles")
    sim.destroy()


def test_empty_simulation(ctx):
    print("[gpu] empty simulation", flush=True)
    sim = MPM(ctx, grid=16, maxParticles=100)
    sim.step()  # mpm.js:584: nothing to do
    g = sim.lattice_gravity()
    assert load_library().mpm_step(sim._h, 4, _ptr(g), 0.4, 0.9) == 0
    assert sim.count == 0 and sim.stepCount == 0
    sim.p2g()
    assert not sim.read_grid().any()
    sim.destroy()


@pytest.mark.parametrize("items,fuse,shape,drift", [(1, 1, 1, 2), (1, 1, 1, 3), (1, 1, 0, 2), (1, 1, 0, 3),
                                                    (1, 0, 1, 3), (0, 0, 1, 2), (0, 1, 1, 2)])
def test_kernel_variants_give_the_same_bits(ctx, items, fuse, shape, drift):
    """G2P by work item with LDS velocity boxes (g2p_items) and G2P fused with the next substep's P2G
    (fuse), on and off, both workgroup shapes of the fused kernel (fuse_shape: 256 threads at 4
    waves per SIMD, 192 at 3) and both P2G tile shapes (drift 2: a 10^3-node window, 3: 12^3): 14
    substeps in one call (3 sorts, 9 fused transitions when both are on) of four materials, then 2
    more calls of 3 substeps (each takes up the P2G the call before carried forward, with fusion on),
    bit-identical to the oracle in every combination (fuse needs g2p_items; (0, 1) runs unfused)."""
    print(f"[gpu] kernel variants g2p_items={items} fuse={fuse} fuse_shape={shape} drift={drift}", flush=True)
    sim = MPM(ctx, grid=40, maxParticles=120000, substeps=8)
    sim.set_tuning(g2p_items=items, fuse=fuse, fuse_shape=shape, drift=drift)
    sim.add({"box": [[0.1, 0.1, 0.1], [0.45, 0.45, 0.45]]}, material="water", velocity=[0.8, -0.5, 0.3], jitter=1.0)
    sim.add({"box": [[0.55, 0.1, 0.1], [0.9, 0.45, 0.45]]}, material="jelly", velocity=[-0.6, 0.2, 0.4], jitter=1.0)
    sim.add({"box": [[0.1, 0.1, 0.55], [0.45, 0.45, 0.9]]}, material="snow", velocity=[0.3, -1.0, -0.2], jitter=1.0)
    sim.add({"box": [[0.55, 0.2, 0.55], [0.9, 0.55, 0.9]]}, material="sand", velocity=[-0.2, -2.0, 0.1], jitter=1.0)
    rec = sim.read_particles()
    sc = _oracle_of(sim, rec)
    g = sim.lattice_gravity()
    L = load_library()
    for k in (14, 3, 3):
        assert L.mpm_step(sim._h, k, _ptr(g), 0.4, 0.9) == 0
        sc.step(k, g)
        _same(sim.read_particles(), sc.read(len(rec)), f"particles after {k} more substeps")
    _same(sim.read_grid(), sc.grid(), "grid")
    sim.destroy()
