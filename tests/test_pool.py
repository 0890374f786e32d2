"""Scene pools (mpm_pool_*): many scenes advanced by shared launches give, scene by scene, exactly
the particles a separate simulation of that scene gives (checked against the oracle, bit for bit and
in the same order), under scene removal and re-admission, with and without captured graphs."""
import numpy as np
import pytest

import oracle_mpm as O
from helpers import bits
from shaderintel_amd import scenes
from shaderintel_amd.mpm import SIError, ScenePool, node_blocks

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


def _scene(seed: int, grid: int, dims=None, gravity=(0.0, -9.81, 0.0)):
    rng = np.random.default_rng(seed)
    blocks = scenes.random_blocks(rng, min(dims) if dims else grid, max_blocks=3, max_edge=14)
    return scenes.host_blocks(grid, blocks, gravity=gravity, dims=dims)


def _oracle(dims, tbl, rec, sort_every=4):
    sc = O.Scene(dims, max(len(rec), 1), 2, sort_every)
    sc.set_materials(tbl)
    sc.write(rec)
    sc.set_count(len(rec))
    return sc


def _same(got, want, what):
    assert got.shape == want.shape, f"{what}: shape {got.shape} vs {want.shape}"
    bad = np.argwhere(bits(got) != bits(want))
    assert len(bad) == 0, f"{what}: {len(bad)} words differ, first at {tuple(bad[0])}"


SPECS = [
    dict(seed=11, grid=32),
    dict(seed=12, grid=48, gravity=(0.0, -9.81, 3.0)),
    dict(seed=13, grid=40, dims=(40, 32, 48)),
    dict(seed=14, grid=64, gravity=(1.0, -4.0, 0.0)),
    dict(seed=15, grid=24),
]


def _build_specs():
    out = []
    for sp in SPECS:
        dims, tbl, rec, g = _scene(sp["seed"], sp["grid"], sp.get("dims"), sp.get("gravity", (0.0, -9.81, 0.0)))
        out.append((dims, tbl, rec, g))
    return out


def test_pool_scenes_match_separate_simulations(ctx):
    """Five scenes of different sizes, shapes, gravities and materials in one pool, 10 substeps:
    each equals the oracle run of that scene alone."""
    _say("[gpu] pool of 5 scenes vs oracle per scene, 10 substeps")
    built = _build_specs()
    total = sum(len(b[2]) for b in built)
    pool = ScenePool(ctx, total + 100, 8, sum(node_blocks(b[0]) for b in built) + 64, 64)
    ids = [pool.add_scene(d, t, r, g) for d, t, r, g in built]
    assert pool.info()["particles"] == total
    pool.step(10)
    for sid, (d, t, r, g) in zip(ids, built):
        sc = _oracle(d, t, r)
        sc.step(10, g)
        _same(pool.read_scene(sid), sc.read(len(r)), f"scene {sid} {d}")
    pool.destroy()


def test_pool_remove_and_readmit(ctx):
    """Remove a scene between steps, admit a new one into the freed slot and blocks: the survivors
    continue exactly, the newcomer starts exactly."""
    _say("[gpu] pool remove / re-admit")
    built = _build_specs()[:3]
    total = sum(len(b[2]) for b in built)
    nb = sum(node_blocks(b[0]) for b in built)
    pool = ScenePool(ctx, total + 20000, 4, nb + 512, 64)
    ids = [pool.add_scene(d, t, r, g) for d, t, r, g in built]
    oracles = [_oracle(d, t, r) for d, t, r, _ in built]
    pool.step(6)
    for sc, (_, _, _, g) in zip(oracles, built):
        sc.step(6, g)
    pool.remove_scene(ids[0])
    with pytest.raises(SIError):
        pool.read_scene(ids[0])
    d2, t2, r2, g2 = _scene(99, 32)
    assert len(r2) <= 20000 + len(built[0][2])
    new_id = pool.add_scene(d2, t2, r2, g2)
    new_oracle = _oracle(d2, t2, r2)
    new_oracle.step_count = 6  # the pool's sort phase (results do not depend on it; particle order does)
    pool.step(7)
    new_oracle.step(7, g2)
    for k in (1, 2):
        oracles[k].step(7, built[k][3])
        _same(pool.read_scene(ids[k]), oracles[k].read(len(built[k][2])), f"survivor {ids[k]}")
    _same(pool.read_scene(new_id), new_oracle.read(len(r2)), "newcomer")
    info = pool.info()
    assert info["scenes"] == 3 and info["particles"] == len(r2) + len(built[1][2]) + len(built[2][2])
    pool.destroy()


def test_pool_graphs_match_eager(ctx):
    """Sort periods replayed from captured graphs give the eager launches' results (several replays,
    both buffer parities, with eager reads in between)."""
    _say("[gpu] pool graphs vs eager, 56 substeps in calls of 8, 13 and 6")
    built = _build_specs()
    total = sum(len(b[2]) for b in built)
    nb = sum(node_blocks(b[0]) for b in built)
    pools = []
    for graphs in (False, True):
        pool = ScenePool(ctx, total, 8, nb, 64)
        pool.use_graphs(graphs)
        ids = [pool.add_scene(d, t, r, g) for d, t, r, g in built]
        pools.append((pool, ids))
    # calls of 8 (two periods), then 13 (three periods and an eager substep) and 6 (an eager period
    # start after the sort phase moved): the eager pool does each sorting substep's P2G ahead of its
    # sort, the graph pool replays plain periods; both give the same bits
    for rep, n in enumerate((8, 8, 8, 13, 6, 13)):
        outs = []
        for pool, ids in pools:
            pool.step(n)
            outs.append([pool.read_scene(i) for i in ids])
        for a, b in zip(*outs):
            _same(b, a, f"call {rep} ({n} substeps)")
    for pool, _ in pools:
        pool.destroy()


def test_pool_admission_limits(ctx):
    """A scene that does not fit (particles, node blocks or slots) is refused with out-of-memory and
    the pool is unchanged."""
    _say("[gpu] pool admission limits")
    d, t, r, g = _scene(21, 32)
    pool = ScenePool(ctx, len(r) + 10, 2, node_blocks(d) + 1, 32)
    a = pool.add_scene(d, t, r, g)
    with pytest.raises(SIError) as e:
        pool.add_scene(d, t, r, g)  # particles
    assert e.value.code == "out-of-memory"
    with pytest.raises(SIError) as e:
        pool.add_scene(d, t, r[:5], g)  # node blocks
    assert e.value.code == "out-of-memory"
    with pytest.raises(SIError) as e:
        pool.add_scene((64, 64, 64), t, r[:5], g)  # larger than max_dim
    assert e.value.code == "invalid-argument"
    assert pool.info() == {"particles": len(r), "scenes": 1, "free_node_blocks": 1}
    pool.step(4)
    sc = _oracle(d, t, r)
    sc.step(4, g)
    _same(pool.read_scene(a), sc.read(len(r)), "after refusals")
    pool.destroy()


def test_template_admission_matches_host_admission(ctx):
    """Scenes admitted from device templates equal scenes admitted from host records."""
    _say("[gpu] template admission vs host admission")
    built = _build_specs()[:3]
    total = sum(len(b[2]) for b in built)
    nb = sum(node_blocks(b[0]) for b in built)
    a = ScenePool(ctx, total, 4, nb, 64)
    b = ScenePool(ctx, total, 4, nb, 64)
    ia = [a.add_scene(d, t, r, g) for d, t, r, g in built]
    tids = [b.add_template(d, t, r, g) for d, t, r, g in built]
    ib = [b.add_from_template(k) for k in tids]
    a.step(6)
    b.step(6)
    for x, y in zip(ia, ib):
        _same(b.read_scene(y), a.read_scene(x), "template scene")
    a.destroy()
    b.destroy()


def test_autotune_is_result_neutral(ctx):
    """The autotuner's trials are real steps with the vmax of the pool's last step: a tuned pool
    equals an untuned pool stepped as far with that vmax (0.01 cells per substep, which binds)."""
    _say("[gpu] autotune result-neutral")
    built = _build_specs()[:2]
    total = sum(len(b[2]) for b in built)
    nb = sum(node_blocks(b[0]) for b in built)
    a = ScenePool(ctx, total, 2, nb, 64)
    b = ScenePool(ctx, total, 2, nb, 64)
    ia = [a.add_scene(d, t, r, g) for d, t, r, g in built]
    ib = [b.add_scene(d, t, r, g) for d, t, r, g in built]
    a.step(4, vmax=0.01)
    b.step(4, vmax=0.01)
    best = a.autotune(4)
    assert 3 <= best["g2p_waves"] <= 6 and 4 <= best["p2g_waves"] <= 6 and best["fuse_shape"] in (0, 1)
    assert best["drift"] in (2, 3)
    b.step(16 * 4, vmax=0.01)
    for x, y in zip(ia, ib):
        _same(a.read_scene(x), b.read_scene(y), "tuned vs untuned")
    a.destroy()
    b.destroy()


def test_server_completes_requests_and_rejects_oversize(ctx):
    """The serving loop admits, steps and retires every request; one that cannot fit is rejected."""
    _say("[gpu] serving loop: 12 requests, drain")
    from shaderintel_amd.mpm import Server
    built = _build_specs()[:3]
    big = max(len(b[2]) for b in built)
    pool = ScenePool(ctx, big + 4000, 8, 2 * max(node_blocks(b[0]) for b in built), 64)
    tids = [pool.add_template(d, t, r, g) for d, t, r, g in built]
    dims, tbl, rec, g = built[0]
    huge = pool.add_template(dims, tbl, np.concatenate([rec] * 20), g)  # more particles than the pool holds
    srv = Server(pool, substeps=4)
    for k in range(12):
        srv.submit(tids[k % 3], 1 + k % 4)
    srv.submit(huge, 2)
    srv.stop(drain=True)
    st = srv.stats()
    assert st["error"] == 0
    assert st["submitted"] == 13 and st["completed"] == 12 and st["rejected"] == 1
    assert st["waiting"] == 0 and st["running"] == 0 and st["resident_particles"] == 0
    assert st["steps"] >= 4 and st["particle_substeps"] > 0
    assert len(srv.latencies("step_ms")) == st["steps"]
    req = srv.latencies("request_s")
    wait = srv.latencies("wait_s")
    assert len(req) == 12 and np.all(req >= wait) and np.all(wait >= 0)
    # scene-step requests: one per step of each completed session (lifetimes 1 + k % 4), each served
    # after it was issued; after a drain nothing is pending, so the censored series is the same
    steps = srv.latencies("scene_step_s")
    assert st["scene_steps"] == len(steps) == sum(1 + k % 4 for k in range(12))
    assert np.all(steps > 0)
    assert np.array_equal(np.sort(srv.latencies("scene_step_age_s")), np.sort(steps))
    srv.destroy()
    assert pool.info()["scenes"] == 0
    pool.destroy()


def test_server_graph_after_keeps_results(ctx):
    """With captured sort periods replayed once the resident scenes are unchanged for 2 steps, served
    scenes equal the same sessions served eagerly (bit for bit)."""
    _say("[gpu] serving loop: graph_after 2 against eager")
    from shaderintel_amd.mpm import Server
    built = _build_specs()[:2]
    out = {}
    for ga in (0, 2):
        pool = ScenePool(ctx, sum(len(b[2]) for b in built) + 1000, 4, 2 * sum(node_blocks(b[0]) for b in built), 64)
        tids = [pool.add_template(d, t, r, g) for d, t, r, g in built]
        srv = Server(pool, substeps=4, graph_after=ga)
        res = [np.zeros((len(built[i][2]), 32), np.float32) for i in range(2)]
        tickets = [srv.submit(tids[0], 9, result=res[0]), srv.submit(tids[1], 6, result=res[1])]
        for t in tickets:
            assert srv.wait(t, 60000)[0] == "done"
        srv.stop(drain=True)
        assert srv.stats()["error"] == 0
        srv.destroy()
        pool.destroy()
        out[ga] = res
    for a, b in zip(out[0], out[2]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
