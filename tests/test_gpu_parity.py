"""GPU parity: the HIP path through the C ABI against the CPU oracle, bit for bit.

The oracle fixes one op order for every float expression; the kernels follow the same order with
no contraction, so particles and grids must match exactly (not just within a tolerance).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_mpm as O
from helpers import bits, randomize_fast
from shaderintel_amd import scenes
from shaderintel_amd.mpm import MPM, SIError

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def _mirror(sim: MPM, rec: np.ndarray, sort_every: int = 4) -> O.Scene:
    sc = O.Scene(sim.dims, max(len(rec), 1), sim.ppcAxis, sort_every)
    sc.set_materials(sim.material_table())
    sc.write(rec)
    sc.set_count(len(rec))
    return sc


def _first_diff(a: np.ndarray, b: np.ndarray) -> str:
    ab, bb = bits(a).reshape(a.shape[0], -1), bits(b).reshape(b.shape[0], -1)
    bad = np.argwhere(ab != bb)
    if len(bad) == 0:
        return "identical"
    i, j = bad[0]
    return f"{len(bad)} words differ; first at row {i} word {j}: {a.reshape(a.shape[0], -1)[i, j]!r} vs {b.reshape(b.shape[0], -1)[i, j]!r}"


def test_add_matches_host_build(ctx):
    _say("[gpu] add() records vs host build, dam64")
    sim = MPM(ctx, **scenes.options("dam64"))
    n = scenes.populate(sim, "dam64")
    assert n == 65536
    dims, tbl, rec, g = scenes.host_build("dam64")
    got = sim.read_particles()
    assert np.array_equal(bits(got), bits(rec)), _first_diff(got, rec)
    assert np.array_equal(bits(sim.material_table()), bits(tbl))
    sim.destroy()


@pytest.mark.parametrize("force_global", [False, True])
def test_p2g_grid_bit_exact(ctx, force_global):
    _say(f"[gpu] p2g bit-exact, dam64 randomized, force_global={force_global}")
    sim = MPM(ctx, **scenes.options("dam64"))
    scenes.populate(sim, "dam64")
    raw = randomize_fast(sim.read_particles(), seed=3)
    sim.write_particles(raw)
    sim.p2g(force_global=force_global)
    got = sim.read_grid()
    sc = _mirror(sim, raw)
    sc.p2g()
    want = sc.grid()
    assert np.array_equal(got, want), _first_diff(got, want)
    assert want[:, 3].sum() > 0
    sim.destroy()


def test_tiled_equals_global_after_steps(ctx):
    """physics.test.js:33-42: 3 frames, _sort, then tiled and global P2G give identical words."""
    _say("[gpu] tiled == global after 3 frames (grid 48)")
    sim = MPM(ctx, grid=48, maxParticles=80000, substeps=4)
    sim.add({"box": [[0.1, 0.1, 0.1], [0.9, 0.4, 0.9]]}, material="water")
    for _ in range(3):
        sim.step()
    sim._sort()
    sim.p2g(force_global=True)
    a = sim.read_grid()
    sim.p2g(force_global=False)
    b = sim.read_grid()
    assert np.array_equal(a, b)
    assert a[:, 3].sum() > 0
    sim.destroy()


def test_shuffle_invariance(ctx):
    _say("[gpu] shuffled particle order gives the same grid")
    sim = MPM(ctx, **scenes.options("dam64"))
    scenes.populate(sim, "dam64")
    raw = randomize_fast(sim.read_particles(), seed=5)
    sim.write_particles(raw)
    sim.p2g()
    a = sim.read_grid()
    perm = np.random.default_rng(11).permutation(len(raw))
    sim.write_particles(raw[perm])
    sim.p2g(force_global=True)
    b = sim.read_grid()
    assert np.array_equal(a, b)
    sim.destroy()


@pytest.mark.parametrize("substeps", [1, 10])
def test_dam_step_bit_exact(ctx, substeps):
    _say(f"[gpu] dam64 {substeps} substeps vs oracle")
    sim = MPM(ctx, **scenes.options("dam64"))
    scenes.populate(sim, "dam64")
    rec = sim.read_particles()
    sc = _mirror(sim, rec)
    g = sim.lattice_gravity()
    from shaderintel_amd.mpm import load_library, _ptr
    L = load_library()
    assert L.mpm_step(sim._h, substeps, _ptr(g), 0.4, 0.9) == 0
    sc.step(substeps, g)
    got, want = sim.read_particles(), sc.read(len(rec))
    assert np.array_equal(bits(got), bits(want)), _first_diff(got, want)
    gg, gw = sim.read_grid(), sc.grid()
    assert np.array_equal(gg, gw), _first_diff(gg, gw)
    sim.destroy()


MATERIAL_SCENES = {
    "water": "water", "jelly": "jelly", "snow": "snow", "sand": scenes.SAND30, "honey": "honey", "rubber": "rubber",
}


@pytest.mark.parametrize("mat", list(MATERIAL_SCENES))
@pytest.mark.parametrize("substeps,sort_every", [(1, 4), (10, 4), (10, 3)])
def test_material_step_bit_exact(ctx, mat, substeps, sort_every):
    """Each material, randomized v/C/tau and a sheared F, 1 and 10 substeps: particles and grid
    identical to the oracle."""
    _say(f"[gpu] material {mat}, {substeps} substeps (sort every {sort_every}) vs oracle")
    sim = MPM(ctx, grid=32, maxParticles=40000, substeps=8, sortEvery=sort_every)
    sim.add({"box": [[0.2, 0.15, 0.2], [0.6, 0.55, 0.6]]}, material=MATERIAL_SCENES[mat], velocity=[0.3, -0.2, 0.1],
            jitter=1.0, noise=0.2)
    raw = sim.read_particles()
    rng = np.random.default_rng(7)
    n = len(raw)
    raw[:, 8:11] = (rng.random((n, 3)) - 0.5) * 0.05
    raw[:, 12:15] = (rng.random((n, 3)) - 0.5) * 0.05
    raw[:, 16:19] = (rng.random((n, 3)) - 0.5) * 0.05
    if MATERIAL_SCENES[mat] not in ("water", "honey"):
        F = np.eye(3)[None] + (rng.random((n, 3, 3)) - 0.5) * 0.3
        raw[:, 20:23], raw[:, 24:27], raw[:, 28:31] = F[:, :, 0], F[:, :, 1], F[:, :, 2]
    sim.write_particles(raw)
    sc = _mirror(sim, raw, sort_every)
    g = sim.lattice_gravity()
    from shaderintel_amd.mpm import load_library, _ptr
    assert load_library().mpm_step(sim._h, substeps, _ptr(g), 0.4, 0.9) == 0
    sc.step(substeps, g)
    got, want = sim.read_particles(), sc.read(n)
    assert np.all(np.isfinite(got))
    assert np.array_equal(bits(got), bits(want)), _first_diff(got, want)
    sim.destroy()


def test_sort_matches_oracle(ctx):
    _say("[gpu] _sort order vs oracle stable sort")
    sim = MPM(ctx, **scenes.options("dam64"))
    scenes.populate(sim, "dam64")
    raw = sim.read_particles()
    perm = np.random.default_rng(2).permutation(len(raw))
    raw = raw[perm]
    sim.write_particles(raw)
    sim._sort()
    got = sim.read_particles()
    sc = _mirror(sim, raw)
    sc.sort()
    want = sc.read(len(raw))
    assert np.array_equal(bits(got), bits(want)), _first_diff(got, want)
    # the reference's invariant (checkSorted on its 4^3-block Morton keys, mpm.js:567-568): block keys
    # are non-decreasing; inside a block the order is (rank in cell, cell numbered x fastest)
    keys = sc.keys(len(raw)).astype(np.int64)
    assert np.all(np.diff(keys >> 6) >= 0)
    blocks = keys >> 6
    m = keys & 63  # the cell's Morton bits x0 y0 z0 x1 y1 z1
    lin = ((m & 1) | ((m >> 2) & 2)) + 4 * (((m >> 1) & 1) | ((m >> 3) & 2)) + 16 * (((m >> 2) & 1) | ((m >> 4) & 2))
    for b in np.unique(blocks)[:50]:
        cells = lin[blocks == b]
        ranks = np.array([np.sum(cells[:i] == c) for i, c in enumerate(cells)])
        order_key = ranks * 64 + cells
        assert np.all(np.diff(order_key) > 0)
    sim.destroy()


SCAN_SIZES = [1, 3, 4, 17, 4096, 4097, 65537, 1 << 20, (1 << 24) + 13]


@pytest.mark.parametrize("n", SCAN_SIZES)
def test_scan(ctx, n):
    """core.test.js:72-110 sizes: exclusive scan against a running sum (u32 wrapping)."""
    _say(f"[gpu] scan n={n}")
    a = np.random.default_rng(n).integers(0, 1 << 20, n, dtype=np.uint32)
    want = np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]).astype(np.uint64) & 0xFFFFFFFF
    got = ctx.scan(a)
    assert np.array_equal(got.astype(np.uint64), want)


@pytest.mark.parametrize("n,nb", [(0, 7), (1, 1), (1000, 1), (65536, 4096), (300001, 262144), (2_000_003, 1 << 24)])
def test_counting_sort(ctx, n, nb):
    """grid.js:70-79 semantics with a stable rank: order is the stable argsort, start the bucket offsets."""
    _say(f"[gpu] counting sort n={n} buckets={nb}")
    keys = np.random.default_rng(n + nb).integers(0, nb, n, dtype=np.uint32)
    order, start = ctx.counting_sort(keys, nb)
    assert np.array_equal(order, np.argsort(keys, kind="stable").astype(np.uint32))
    assert start[nb] == n
    assert np.array_equal(start, np.searchsorted(np.sort(keys), np.arange(nb + 1)).astype(np.uint32))


def test_conservation_law(ctx):
    """verify/index.js:322-329 (law mpm-conservation) with checkMPM's tolerance (:175-193)."""
    _say("[gpu] mpm-conservation law")
    sim = MPM(ctx, grid=32, maxParticles=50000)
    sim.add({"box": [[0.2, 0.2, 0.2], [0.6, 0.6, 0.6]]}, material="water", velocity=[0.3, 0.1, -0.2])
    for _ in range(10):
        sim.step()
    raw = sim.read_particles()
    sim.p2g()
    g = sim.read_grid().astype(np.float64)
    FP = 2.0 ** 22
    m = 1 / 8
    n = len(raw)
    pm = m * n
    pv = m * raw[:, 4:7].astype(np.float64).sum(0)
    gm = g[:, 3].sum() / FP
    gp = g[:, :3].sum(0) / FP
    mass_err = abs(gm - pm) / pm
    mom_err = np.linalg.norm(gp - pv) / pm
    assert mass_err < 1e-4 and mom_err < 1e-4 + n * 27 * 4 / FP / pm, (mass_err, mom_err)
    sim.destroy()


def test_errors(ctx):
    sim = MPM(ctx, grid=16, maxParticles=1024)
    rec = np.zeros((1, 32), np.float32)
    rec[0, :3] = [100, 5, 5]
    with pytest.raises(SIError) as e:
        sim.write_particles(rec)
    assert e.value.code == "invalid-argument"
    with pytest.raises(SIError) as e:
        sim.add({"box": [[0.1, 0.1, 0.1], [0.2, 0.2, 0.2]]}, material={"kind": "fluid", "density": 1000})
    assert e.value.code == "invalid-argument"
    sim.destroy()
    sim.destroy()
    with pytest.raises(SIError):
        sim.step()


@pytest.mark.parametrize("drift", [2, 3])
def test_fast_particles_use_fallback_and_stay_exact(ctx, drift):
    """Particles fast enough to leave their block's window between sorts (sortEvery 8, |v| near the
    CFL clamp) take the global-atomic fallback; the result stays identical to the oracle with either
    window (drift 2: a 10^3-node tile, 3: 12^3)."""
    _say(f"[gpu] fallback path: fast particles, sortEvery 8, drift {drift}")
    sim = MPM(ctx, grid=48, maxParticles=60000, substeps=8, sortEvery=8, gravity=[0, 0, 0])
    sim.set_tuning(drift=drift)
    sim.add({"box": [[0.2, 0.2, 0.2], [0.45, 0.45, 0.45]]}, material="water", velocity=[20.0, 15.0, -10.0], jitter=1.0)
    sim.add({"box": [[0.55, 0.5, 0.5], [0.8, 0.75, 0.75]]}, material="jelly", velocity=[-25.0, -5.0, 12.0], jitter=1.0)
    rec = sim.read_particles()
    sc = _mirror(sim, rec, sort_every=8)
    g = sim.lattice_gravity()
    from shaderintel_amd.mpm import load_library, _ptr
    assert load_library().mpm_step(sim._h, 24, _ptr(g), 0.4, 0.9) == 0
    sc.step(24, g)
    got, want = sim.read_particles(), sc.read(len(rec))
    assert np.array_equal(bits(got), bits(want)), _first_diff(got, want)
    gg, gw = sim.read_grid(), sc.grid()
    assert np.array_equal(gg, gw), _first_diff(gg, gw)
    sim.destroy()


def test_p2g_after_write_without_sort(ctx):
    """p2g on freshly written, unsorted particles bins them through a permutation without reordering
    them: the read-back order is unchanged and the grid equals the oracle's."""
    _say("[gpu] p2g on unsorted particles keeps their order")
    sim = MPM(ctx, grid=32, maxParticles=20000)
    sim.add({"box": [[0.2, 0.2, 0.2], [0.6, 0.5, 0.6]]}, material="snow")
    raw = randomize_fast(sim.read_particles(), seed=9)
    raw = raw[np.random.default_rng(4).permutation(len(raw))]
    sim.write_particles(raw)
    sim.p2g()
    assert np.array_equal(bits(sim.read_particles()), bits(raw))
    sc = _mirror(sim, raw)
    sc.p2g()
    assert np.array_equal(sim.read_grid(), sc.grid())
    sim.destroy()


@pytest.mark.parametrize("sort_every", [4, 3])
def test_split_calls_match_one_run(ctx, sort_every):
    """Substeps split over calls of 1-5 (calls that start on a sort, which open with the sort's P2G
    over the previous work items, mid-period and right after one) give the oracle's bits of one
    uninterrupted run, particles and the grid after every call (also after a call whose last substep
    sorts, whose grid update ran before the sort); so does a call that starts on a sort after the
    particles were rewritten (the write drops the work items, so that call sorts first)."""
    _say(f"[gpu] split calls vs one oracle run (sort every {sort_every})")
    sim = MPM(ctx, grid=32, maxParticles=40000, substeps=8, sortEvery=sort_every)
    sim.add({"box": [[0.15, 0.15, 0.15], [0.45, 0.5, 0.45]]}, material="snow", velocity=[0.3, -0.2, 0.1], noise=0.2)
    sim.add({"box": [[0.55, 0.15, 0.55], [0.85, 0.45, 0.85]]}, material=scenes.SAND30, velocity=[-0.2, 0.1, 0.0])
    raw = sim.read_particles()
    sc = _mirror(sim, raw, sort_every)
    g = sim.lattice_gravity()
    from shaderintel_amd.mpm import load_library, _ptr
    L = load_library()
    n = len(raw)
    for calls in ([1, 3, 4, 5, 2, 1], [4, 4, 8]):
        for k in calls:
            assert L.mpm_step(sim._h, k, _ptr(g), 0.4, 0.9) == 0
            sc.step(k, g)
            # the grid after each call: a call whose last substep sorts ran that substep's grid update
            # before the sort, over the previous work items' S
            gg, gw = sim.read_grid(), sc.grid()
            assert np.array_equal(gg, gw), f"grid after a call of {k} in {calls}: {_first_diff(gg, gw)}"
        got, want = sim.read_particles(), sc.read(n)
        assert np.array_equal(bits(got), bits(want)), f"calls {calls}: {_first_diff(got, want)}"
    # rewrite (same order): shift every particle by up to 3 cells, then step from a sort
    moved = got.copy()
    moved[:, 0:3] += (np.random.default_rng(5).random((n, 3)) - 0.5) * 6.0
    moved[:, 0:3] = np.clip(moved[:, 0:3], 3.0, 28.0)
    sim.write_particles(moved)
    sc.write(moved)
    assert L.mpm_step(sim._h, 2 * sort_every, _ptr(g), 0.4, 0.9) == 0
    sc.step(2 * sort_every, g)
    got, want = sim.read_particles(), sc.read(n)
    assert np.array_equal(bits(got), bits(want)), f"after rewrite: {_first_diff(got, want)}"
    gg, gw = sim.read_grid(), sc.grid()
    assert np.array_equal(gg, gw), _first_diff(gg, gw)
    sim.destroy()


def _three_materials(ctx, grid=32):
    sim = MPM(ctx, grid=grid, maxParticles=60000, substeps=8, sortEvery=4)
    sim.add({"box": [[0.12, 0.15, 0.12], [0.42, 0.45, 0.42]]}, material="snow", velocity=[0.3, -0.2, 0.1], noise=0.2)
    sim.add({"box": [[0.55, 0.15, 0.55], [0.85, 0.4, 0.85]]}, material=scenes.SAND30, velocity=[-0.2, 0.1, 0.0])
    sim.add({"box": [[0.15, 0.55, 0.55], [0.45, 0.8, 0.85]]}, material="jelly", velocity=[0.0, -0.3, 0.0])
    return sim


@pytest.mark.parametrize("carry", [1, 0])
def test_frames_with_reads_between_match_one_run(ctx, carry):
    """The JavaScript host's loop: frames of 8 substeps (MPM.step, which uploads the colliders first)
    with particle and grid reads between them. With MPM_TUNE_CARRY on, a call that follows another
    with only reads (and collider uploads) between them ends with its G2P
    fused with the next substep's P2G, and the next call takes that P2G up (the launch counts show
    it); particles and grid (the last grid update's velocities) equal the oracle's after every call,
    with the carry on and off."""
    _say(f"[gpu] frames with reads between them vs oracle (carry {carry})")
    sim = _three_materials(ctx)
    sim.setColliders([{"type": "sphere", "center": [0.5, 0.2, 0.5], "radius": 0.1, "friction": 0.4,
                       "velocity": [0.1, 0.0, 0.0]}])
    sim.set_tuning(carry=carry)
    raw = sim.read_particles()
    sc = _mirror(sim, raw)
    sc.set_colliders(sim.collider_table(), 1)
    g = sim.lattice_gravity()
    n = len(raw)
    sim.set_profiling(2)
    for frame in range(4):
        sim.step()  # uploads the collider table (kept: colliders act in the grid update), 8 substeps
        sc.step(8, g)
        got, want = sim.read_particles(), sc.read(n)
        assert np.array_equal(bits(got), bits(want)), f"frame {frame}: {_first_diff(got, want)}"
        gg, gw = sim.read_grid(), sc.grid()
        assert np.array_equal(gg, gw), f"frame {frame} grid: {_first_diff(gg, gw)}"
    kt = sim.kernel_times()
    # the first call after set_profiling ends with a plain G2P; with the carry, the next three end
    # fused and open without a P2G of their own
    assert (kt["g2p"][1], kt["g2p_p2g"][1]) == ((1, 31) if carry else (4, 28)), kt
    sim.destroy()


def test_dropped_carried_p2g_leaves_no_sums_in_a(ctx):
    """Particles moving together at 0.88 cells per substep leave their work items' windows (drift 2) within a sort
    period, so the P2G a call carries forward at a sort boundary sends many of them through A, also
    into blocks of S (which it stamps but does not list in D). A change between the calls drops that
    P2G; the next call's P2G must find none of its sums left in A (an earlier version cleared only
    the D blocks, and the next grid update over a restamped block counted them twice): the next
    call's grid bookkeeping (mass, momentum) and particles equal the oracle's."""
    _say("[gpu] dropped carried P2G with fast particles vs oracle")
    sim = _three_materials(ctx)
    sim.set_tuning(drift=2)
    raw = sim.read_particles()
    n = len(raw)
    raw[:, 4:7] = np.array([0.8, -0.3, 0.2], np.float32)  # one velocity: the grid keeps it
    sim.write_particles(raw)
    sc = _mirror(sim, raw)
    g = sim.lattice_gravity()
    from shaderintel_amd.mpm import load_library, _ptr
    L = load_library()
    for _ in range(2):  # the second call carries its last P2G (the next sort's) forward
        assert L.mpm_step(sim._h, 4, _ptr(g), 0.4, 0.9) == 0
        sc.step(4, g)
    for label, change in (("a knob set", lambda: sim.set_tuning(carry=1)), ("a drift change", lambda: sim.set_tuning(drift=3))):
        sim.set_grid_law(True)
        sc.set_grid_law(True)
        change()  # drops the carried P2G
        for _ in range(2):
            assert L.mpm_step(sim._h, 4, _ptr(g), 0.4, 0.9) == 0
            sc.step(4, g)
        assert sim.read_grid_law()["raw"] == [int(x) for x in sc.read_grid_law()[:14].astype(np.int64)], label
        got, want = sim.read_particles(), sc.read(n)
        assert np.array_equal(bits(got), bits(want)), f"after {label}: {_first_diff(got, want)}"
    sim.destroy()


def test_changes_between_calls_discard_the_carried_p2g(ctx):
    """Between calls that carry a P2G forward: a rewrite of the particles, a new material table, a
    retime, a new step count, an explicit sort and an explicit P2G. Each discards the carried P2G
    (otherwise the next call would start from sums of the old state); after each, the next call's
    particles equal the oracle's doing the same."""
    _say("[gpu] changes between calls vs oracle")
    sim = _three_materials(ctx)
    raw = sim.read_particles()
    sc = _mirror(sim, raw)
    g = sim.lattice_gravity()
    from shaderintel_amd.mpm import load_library, _ptr
    L = load_library()
    n = len(raw)
    rng = np.random.default_rng(3)

    def frame(label):
        for _ in range(2):  # the second call carries its last P2G forward
            assert L.mpm_step(sim._h, 8, _ptr(g), 0.4, 0.9) == 0
            sc.step(8, g)
        got, want = sim.read_particles(), sc.read(n)
        assert np.array_equal(bits(got), bits(want)), f"after {label}: {_first_diff(got, want)}"
        return got

    cur = frame("start")
    moved = cur.copy()
    moved[:, 4:7] += (rng.random((n, 3)) - 0.5).astype(np.float32) * 0.2  # velocities: P2G's momentum
    sim.write_particles(moved)
    sc.write(moved)
    frame("a rewrite")
    tbl = sim.material_table()
    tbl[:3, 1] *= 1.25  # particle masses: P2G's mass and momentum
    assert L.mpm_set_materials(sim._h, _ptr(tbl), 8) == 0
    sc.set_materials(tbl)
    frame("a new material table")
    assert L.mpm_rescale(sim._h, 0.8) == 0
    sc.rescale(0.8)
    frame("a retime")
    sim.stepCount = sim.stepCount + 3
    sc.step_count = sc.step_count + 3
    frame("a new step count")
    sim._sort()
    sc.sort()
    frame("a sort")
    sim.p2g()
    sc.p2g()
    frame("an explicit P2G")
    sim.destroy()


@pytest.mark.parametrize("seed,fast", [(1, False), (2, False), (3, True)])
def test_random_call_sequences_match_the_oracle(ctx, seed, fast):
    """A seeded sequence of 40 calls mixing steps of 1-9 substeps with everything a caller may do
    between them: particle and grid reads, particle rewrites, a new material table, a retime, a new
    step count or sort period, an explicit sort or P2G, a substep split at its grid update with the
    grid edited in between, and result-neutral kernel variants (window, workgroup shape, fusion,
    items, the carried P2G). Particles equal the oracle's after every call,
    and the grid after every read. fast: every particle starts at 0.88 cells per substep in one
    direction, so many leave their work items' windows within a sort period (P2G's fallback)."""
    _say(f"[gpu] random call sequence {seed} vs oracle" + (" (fast particles)" if fast else ""))
    sim = _three_materials(ctx)
    raw = sim.read_particles()
    if fast:
        raw[:, 4:7] = np.array([0.8, -0.3, 0.2], np.float32)
        sim.write_particles(raw)
    sc = _mirror(sim, raw)
    g = sim.lattice_gravity()
    from shaderintel_amd.mpm import load_library, _ptr
    L = load_library()
    n = len(raw)
    rng = np.random.default_rng(seed)
    ops = ["step"] * 10 + ["grid", "write", "mats", "retime", "count", "every", "sort", "p2g", "tune", "tune", "split"]
    log = []
    every = [4]
    for _ in range(40):
        op = ops[rng.integers(len(ops))]
        if op == "step":
            k = int(rng.integers(1, 10))
            assert L.mpm_step(sim._h, k, _ptr(g), 0.4, 0.9) == 0
            sc.step(k, g)
            log.append(f"step {k}")
        elif op == "grid":
            gg, gw = sim.read_grid(), sc.grid()
            assert np.array_equal(gg, gw), f"grid after {log[-4:]}: {_first_diff(gg, gw)}"
            log.append("grid")
        elif op == "write":
            rec = sim.read_particles()
            rec[:, 4:7] += ((rng.random((n, 3)) - 0.5) * 0.1).astype(np.float32)
            sim.write_particles(rec)
            sc.write(rec)
            log.append("write")
        elif op == "mats":
            tbl = sim.material_table()
            tbl[:3, 1] *= np.float32(rng.uniform(0.8, 1.25))
            assert L.mpm_set_materials(sim._h, _ptr(tbl), 8) == 0
            sc.set_materials(tbl)
            log.append("mats")
        elif op == "retime":
            k = float(np.float32(rng.uniform(0.8, 1.2)))
            assert L.mpm_rescale(sim._h, k) == 0
            sc.rescale(k)
            log.append("retime")
        elif op == "count":
            c = sim.stepCount + int(rng.integers(1, 4))
            sim.stepCount = c
            sc.step_count = c
            log.append(f"stepCount {c}")
        elif op == "every":
            e = int(rng.integers(1, 6))
            assert L.mpm_set_sort_every(sim._h, e) == 0
            sc.set_sort_every(e)
            every[0] = e
            log.append(f"sortEvery {e}")
        elif op == "sort":
            sim._sort()
            sc.sort()
            log.append("sort")
        elif op == "split":  # a substep split at its grid update, the grid edited in between
            sim.substep_begin()
            words = sim.read_grid()
            if sc.step_count % every[0] == 0:
                sc.sort()
            sc.p2g()
            sc.grid_update(g, 0.4)
            assert np.array_equal(words, sc.grid()), f"split-substep grid after {log[-4:]}"
            v = words.view(np.float32).reshape(-1, 4).copy()
            v[::3, :3] *= np.float32(0.5)
            edited = v.view(np.int32)
            sim.write_grid(edited)
            sim.substep_end()
            sc.write_grid(edited)
            sc.g2p(0.9)
            sc.step_count = sc.step_count + 1
            log.append("split")
        elif op == "p2g":
            sim.p2g()
            sc.p2g()
            gg, gw = sim.read_grid(), sc.grid()
            assert np.array_equal(gg, gw), f"P2G grid after {log[-4:]}: {_first_diff(gg, gw)}"
            log.append("p2g")
        else:
            knob = ["drift", "fuse_shape", "fuse", "g2p_items", "carry"][rng.integers(5)]
            v = int(rng.integers(2, 4)) if knob == "drift" else int(rng.integers(0, 2))
            sim.set_tuning(**{knob: v})
            log.append(f"{knob}={v}")
        got, want = sim.read_particles(), sc.read(n)
        assert np.array_equal(bits(got), bits(want)), f"after {log[-4:]}: {_first_diff(got, want)}"
    sim.destroy()
