#!/usr/bin/env python3
"""Benchmark: MLS-MPM particle-substeps/s on one MI355X.

A "step" is one MLS-MPM substep (sort every 4th, P2G, grid update, G2P) over the whole scene. The
default workload is the 256^3-grid, 8,388,608-particle four-material scene (water, Neo-Hookean
jelly, snow with SVD plasticity, Drucker-Prager sand), fp32, built through the MPM world API with
seeded inputs. Particles and grid stay resident in HBM; the timed region is exactly --steps
substeps between two stream synchronizations. Inside it the dominant kernel (G2P) is timed by its
own dispatches (hipExtLaunchKernelGGL start/stop events: no packet is added to the stream). A second
pass of --steps substeps with HIP events between the phases gives the per-phase breakdown; those
events add packets, so that pass is reported beside the throughput, not as it.

The dominant kernel's HBM figure is also set against a device-to-device copy measured in the same
run. "small_configs" reports the 64^3 dam break (65,536 particles, 100 substeps) and the 128^3
water + jelly scene (1,048,576 particles, 1000 substeps), best of 3 runs each (skip with
--no-small); both are cache- or launch-bound. The CPU baseline is the oracle on 10 substeps of
the bench scene, plus the dam break in full and 10 substeps of the 128^3 scene.

Two more sections run by default (skip with --no-batched / --no-serve):
- "batched": 64 independent scenes (edges 32-128, 1e4-1e6 particles, one material box each, master
  seed 1) stepped as one scene pool (one launch per phase for all scenes, with and without captured
  sort periods, autotuned shapes) against the same scenes stepped per scene: as separate
  simulations in turn (the reference's one-object-per-scene pattern on a single queue), on 64
  concurrent streams, and as one-scene pools replaying captured sort periods on 64 streams.
- "serve": the native serving loop under scene-step requests at --serve-rate requests/s for
  --serve-seconds (plus one overload rate, --serve-overload): sessions drawn from those scenes
  arrive as a Poisson stream at rate / 300 and live a geometric number of 4-substep steps (mean
  300), each step one request issued when the session's previous step is done; admitted FIFO into
  a pool sized to a 200 GB HBM budget. Reports p50/p99 device step time, p50/p99 request latency
  over the served requests and censored (pending requests counted by their age at the end of the
  window), and the served request rate.

The roofline record follows the contract: "achieved" and "frac" count the algorithmic bytes (384 B
per particle-substep) over the dominant kernel's average launch time. Beside them, "traffic" is the
HBM bytes one launch moves (committed rocprofv3 counter passes, scaled by a copy of known size),
"traffic_frac" that traffic's rate against the 8 TB/s peak, and "limiter" what the committed SQ
counters show bounds the kernel.

Prints one JSON line (the last line of stdout).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MPM particle-substeps/s at 8M particles (1 GPU); batched-serve p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle for each kernel: P2G reads the 128-byte record; G2P reads and
# writes it (256 bytes); the fused kernel does one G2P and the next substep's P2G (384 bytes, the
# whole particle-substep). Grid and sort traffic are not credited.
ALGO_BYTES = {"mpm p2g": 128, "mpm g2p": 256, "mpm g2p_p2g": 384}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_rate(scene: str, substeps: int) -> tuple:
    """The oracle (CPU restatement of the reference algorithm) on the full scene for `substeps`
    substeps, all host threads: (particle-substeps/s, particles, threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_mpm as O
    from shaderintel_amd import scenes

    dims, tbl, rec, g = scenes.host_build(scene)
    sc = O.Scene(dims, len(rec), 2, 4)
    sc.set_materials(tbl)
    sc.write(rec)
    sc.set_count(len(rec))
    del rec
    threads = O.lib().or_num_threads()
    t0 = time.perf_counter()
    sc.step(substeps, g)
    dt = time.perf_counter() - t0
    n = scenes.expected_count(scene)
    return n * substeps / dt, n, threads


def cpu_baseline(scene: str, substeps: int, small: bool = True) -> dict:
    """The oracle on a bounded sample of the bench workload (the full scene for `substeps` substeps),
    and, with `small`, on the 64^3 dam break for its full 100 substeps and the 128^3 water + jelly
    scene for 10 substeps."""
    rate, n, threads = cpu_rate(scene, substeps)
    sorts = (substeps + 3) // 4
    out = {"value": rate, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
           "sample": f"{scene}: all {n} particles, {substeps} substeps ({sorts} sorts), OpenMP over {threads} "
                     "threads; CPU restatement of the reference algorithm, not the reference's WebGPU path"}
    if small:
        out["small_configs"] = {}
        for name, k in (("dam64", 100), ("mixed128", 10)):
            r, m, _ = cpu_rate(name, k)
            out["small_configs"][name] = {"particles": m, "substeps": k, "particle_substeps_per_s": round(r, 1)}
    return out


def gpu_small_configs(ctx) -> dict:
    """The 64^3 dam break (65,536 particles, 100 substeps) and the 128^3 water + jelly scene
    (1,048,576 particles, 1000 substeps) on the GPU: best of 3 batches, each the scene's full run,
    between stream synchronizations. Both are cache- or launch-bound, reported in absolute terms."""
    from shaderintel_amd import scenes
    from shaderintel_amd.mpm import MPM, load_library, _ptr, _check
    L = load_library()
    out = {}
    for name, k in (("dam64", 100), ("mixed128", 1000)):
        sim = MPM(ctx, **scenes.options(name))
        n = scenes.populate(sim, name)
        g = sim.lattice_gravity()
        _check(L.mpm_step(sim._h, 8, _ptr(g), 0.4, 0.9), "step")
        sim.sync()
        best = None
        for _ in range(3):
            t0 = time.perf_counter()
            _check(L.mpm_step(sim._h, k, _ptr(g), 0.4, 0.9), "step")
            sim.sync()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        sim.destroy()
        out[name] = {"particles": n, "substeps": k, "ms_per_substep": round(best * 1e3 / k, 4),
                     "particle_substeps_per_s": round(n * k / best, 1)}
        log(f"{name}: {n} particles, {k} substeps, best {best * 1e3:.1f} ms")
    return out


def copy_bandwidth(nbytes: int = 1 << 31, reps: int = 5) -> float:
    """Device-to-device copy rate (GB/s, bytes read + written) of the HIP runtime's copy kernel on
    an `nbytes` buffer, timed with HIP events on the null stream: the measured ceiling the path's
    HBM figure is set against, beside the 8 TB/s spec peak."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    try:
        assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
        assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # warm-up, hipMemcpyDeviceToDevice
        best = None
        for _ in range(reps):
            assert hip.hipEventRecord(e0, None) == 0
            assert hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None) == 0
            assert hip.hipEventRecord(e1, None) == 0
            assert hip.hipEventSynchronize(e1) == 0
            ms = ctypes.c_float()
            assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
            best = ms.value if best is None else min(best, ms.value)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def copy_calibration(nbytes: int) -> None:
    """Three device-to-device copies of nbytes (the HIP runtime's copy kernel): a dispatch whose HBM
    traffic is known (nbytes read, nbytes written), to scale byte counters by."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 copies of {nbytes} bytes (each reads and writes {nbytes} bytes)")


def pctl(a, q):
    return round(float(np.percentile(a, q)), 3) if len(a) else None


_BUILT = {}


def batched_scenes(seed: int = 1, count: int = 64):
    """The batched workload's scenes built on the host (cached: the serving runs reuse them)."""
    from shaderintel_amd import scenes
    if (seed, count) not in _BUILT:
        _BUILT[(seed, count)] = [scenes.host_blocks(edge, blocks) for edge, blocks in scenes.batched_blocks(seed, count)]
    return _BUILT[(seed, count)]


def batched(ctx, substeps: int, seed: int = 1, count: int = 64) -> dict:
    """The batched workload (64 scenes from one master seed, see scenes.batched_blocks): one scene
    pool stepped with one launch per phase (captured sort periods, autotuned shapes), and without
    captured graphs, against the same scenes stepped per scene: (a) as separate simulations in turn
    on one queue (the reference's pattern), (b) as separate simulations on their own streams, all
    enqueued before one wait, and (c) as one-scene pools with captured sort periods on their own
    streams. Each figure is the best of two timed runs."""
    from shaderintel_amd.mpm import MPM, ScenePool, load_library, node_blocks, _ptr, _check
    L = load_library()
    built = batched_scenes(seed, count)
    total = sum(len(b[2]) for b in built)
    log(f"batched: {count} scenes, {total} particles")

    def best(fn, reps=2):
        out = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            out.append(time.perf_counter() - t0)
        return min(out)

    pool = ScenePool(ctx, total, count, sum(node_blocks(b[0]) for b in built), max(b[0][0] for b in built))
    for d, t, r, g in built:
        pool.add_scene(d, t, r, g)
    pool.use_graphs(True)
    tuned = pool.autotune(4)
    pool.step(8)
    pool.sync()

    def run_pool():
        pool.step(substeps)
        pool.sync()
    dt_pool = best(run_pool)
    pool.use_graphs(False)
    pool.step(8)
    pool.sync()
    dt_pool_eager = best(run_pool)
    pool.destroy()
    log(f"batched pool: {substeps} substeps in {dt_pool * 1e3:.1f} ms with graphs, {dt_pool_eager * 1e3:.1f} ms eager "
        f"(tuned {tuned})")
    sims = []
    for d, t, r, g in built:
        sim = MPM(ctx, grid=d[0], size=d[0] / 64.0, maxParticles=len(r), substeps=8)
        _check(L.mpm_set_materials(sim._h, _ptr(np.ascontiguousarray(t)), 8), "materials")
        sim.write_particles(r)
        _check(L.mpm_set_count(sim._h, len(r)), "count")
        sims.append((sim, np.ascontiguousarray(g)))
    for sim, g in sims:
        _check(L.mpm_step(sim._h, 8, _ptr(g), 0.4, 0.9), "step")
    for sim, _ in sims:
        sim.sync()

    def in_turn():  # (a) one simulation object per scene, each stepped in turn on one queue
        for sim, g in sims:
            _check(L.mpm_step(sim._h, substeps, _ptr(g), 0.4, 0.9), "step")
            sim.sync()

    def on_streams():  # (b) the same simulations on their own HIP streams (they overlap)
        for sim, g in sims:
            _check(L.mpm_step(sim._h, substeps, _ptr(g), 0.4, 0.9), "step")
        for sim, _ in sims:
            sim.sync()
    dt_seq = best(in_turn)
    dt_sep = best(on_streams)
    for sim, _ in sims:
        sim.destroy()
    pools = []
    for d, t, r, g in built:
        p1 = ScenePool(ctx, len(r), 1, node_blocks(d), d[0])
        p1.add_scene(d, t, r, g)
        p1.use_graphs(True)
        p1.step(8)
        pools.append(p1)
    for p1 in pools:
        p1.sync()

    def graphs_on_streams():  # (c) one-scene pools replaying captured sort periods, own streams
        for p1 in pools:
            p1.step(substeps)
        for p1 in pools:
            p1.sync()
    dt_gr = best(graphs_on_streams)
    for p1 in pools:
        p1.destroy()
    log(f"batched separate: {substeps} substeps in {dt_seq * 1e3:.1f} ms in turn, {dt_sep * 1e3:.1f} ms on 64 streams, "
        f"{dt_gr * 1e3:.1f} ms as graphs on 64 streams")
    rate = {k: total * substeps / dt for k, dt in
            (("pool", dt_pool), ("pool_eager", dt_pool_eager), ("seq", dt_seq), ("sep", dt_sep), ("gr", dt_gr))}
    return {"workload": f"{count} scenes (master seed {seed}): edges 32-128, 1e4-1e6 particles each, "
                        f"water/jelly/snow/sand, {total} particles, {substeps} substeps",
            "particles": total, "substeps": substeps,
            "pool_particle_substeps_per_s": round(rate["pool"], 1),
            "pool_eager_particle_substeps_per_s": round(rate["pool_eager"], 1),
            "per_scene_in_turn_particle_substeps_per_s": round(rate["seq"], 1),
            "per_scene_streams_particle_substeps_per_s": round(rate["sep"], 1),
            "per_scene_graphs_streams_particle_substeps_per_s": round(rate["gr"], 1),
            "speedup_vs_per_scene_in_turn": round(rate["pool"] / rate["seq"], 2),
            "speedup_vs_per_scene_streams": round(rate["pool"] / rate["sep"], 2),
            "speedup_vs_per_scene_graphs_streams": round(rate["pool"] / rate["gr"], 2),
            "tuned": tuned, "pool_ms": round(dt_pool * 1e3, 2), "pool_eager_ms": round(dt_pool_eager * 1e3, 2),
            "in_turn_ms": round(dt_seq * 1e3, 2), "streams_ms": round(dt_sep * 1e3, 2),
            "graphs_streams_ms": round(dt_gr * 1e3, 2)}


def serve(ctx, rate: float, seconds: float, budget_gb: float, substeps: int = 4, lifetime_mean: float = 300,
          seed: int = 2, graph_after: int = 8) -> dict:
    """Serving: a stream of scene-step requests at `rate` requests/s for `seconds`. A request advances
    one scene by one step of `substeps` substeps. Scenes (sessions) are drawn from the 64
    batched-workload scenes and live a geometric number of steps (mean `lifetime_mean`); a session
    issues its next step request when its previous step is done, so sessions arrive as a Poisson
    stream at rate / lifetime_mean and the step requests come at `rate` on average. One pool sized to
    the HBM budget; the native serving loop admits, steps every resident session together and retires
    (continuous batching). Latencies: per scene-step request (issue -> its step done), over the served
    ones and censored (pending requests by their age at the end of the window)."""
    from shaderintel_amd.mpm import ScenePool, Server
    built = batched_scenes(1, 64)
    tmpl_bytes = sum(len(b[2]) for b in built) * 128
    budget = budget_gb * 1e9 - tmpl_bytes
    max_scenes, max_dim = 2048, 128

    def cfg(cap):  # node blocks: ~1 per 20 particles for this scene mix (grids are sparse)
        return cap, max_scenes, max(cap // 20, 32768), max_dim

    lo, hi = 1 << 20, 1 << 30
    while hi - lo > (1 << 20):
        mid = (lo + hi) // 2
        lo, hi = (mid, hi) if ScenePool.bytes_for(*cfg(mid)) <= budget else (lo, mid)
    pool = ScenePool(ctx, *cfg(lo))
    session_rate = rate / lifetime_mean
    log(f"serve: pool of {lo} particles / {cfg(lo)[2]} node blocks = {ScenePool.bytes_for(*cfg(lo)) / 1e9:.1f} GB "
        f"(+{tmpl_bytes / 1e9:.1f} GB templates) for a {budget_gb} GB budget; {rate} step req/s "
        f"({session_rate:.2f} sessions/s) for {seconds} s")
    tids = [pool.add_template(d, t, r, g) for d, t, r, g in built]
    rng = np.random.default_rng(seed)
    srv = Server(pool, substeps=substeps, graph_after=graph_after)
    t0 = time.perf_counter()
    t_next = 0.0
    while True:
        t_next += rng.exponential(1.0 / session_rate)
        if t_next > seconds:
            break
        delay = t0 + t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        srv.submit(tids[int(rng.integers(len(tids)))], int(rng.geometric(1.0 / lifetime_mean)))
    delay = t0 + seconds - time.perf_counter()
    if delay > 0:
        time.sleep(delay)
    wall = time.perf_counter() - t0
    ages = srv.latencies("scene_step_age_s")  # taken at the end of the window, before the stop
    served = srv.latencies("scene_step_s")
    srv.stop(drain=False)
    st = srv.stats()
    steps, req = srv.latencies("step_ms"), srv.latencies("request_s")
    srv.destroy()
    pool.destroy()
    log(f"serve: {st}")
    mean_particles = float(np.mean([len(b[2]) for b in built]))
    return {"workload": f"scene-step requests at {rate}/s for {seconds} s: sessions drawn from the 64-scene batched "
                        f"mix arrive as a Poisson stream at {session_rate:.3g}/s and live a geometric number of "
                        f"steps (mean {lifetime_mean}), one request per step of {substeps} substeps, issued when "
                        f"the session's previous step is done; {budget_gb} GB HBM budget"
                        + (f"; captured sort periods once the resident scenes are unchanged for {graph_after} steps"
                           if graph_after else "; eager launches"),
            "rate_req_s": rate, "sessions_submitted": st["submitted"], "sessions_completed": st["completed"],
            "sessions_rejected": st["rejected"], "sessions_waiting_at_end": st["waiting"],
            "sessions_resident_at_end": st["running"], "resident_particles_at_end": st["resident_particles"],
            "steps": st["steps"], "scene_steps_served": st["scene_steps"],
            "scene_steps_per_s": round(st["scene_steps"] / wall, 1),
            "step_latency_ms_p50": pctl(steps, 50), "step_latency_ms_p99": pctl(steps, 99),
            "request_latency_ms_p50": pctl(served * 1e3, 50), "request_latency_ms_p99": pctl(served * 1e3, 99),
            "request_latency_censored_ms_p50": pctl(ages * 1e3, 50),
            "request_latency_censored_ms_p99": pctl(ages * 1e3, 99),
            "pending_at_end": int(len(ages) - len(served)),
            "session_latency_s_p50": pctl(req, 50), "session_latency_s_p99": pctl(req, 99), "graph_after": graph_after,
            "particle_substeps_per_s": round(st["particle_substeps"] / wall, 1),
            "mean_particles_per_request": round(mean_particles, 1), "error": st["error"]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scene", default="mixed256")
    ap.add_argument("--cpu-substeps", type=int, default=10)
    ap.add_argument("--time-every", type=int, default=7,
                    help="time every N-th G2P launch of the timed region with its own dispatch events (1: all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep", action="store_true", help="time each tuning knob setting and exit")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-scene batched comparison")
    ap.add_argument("--batched-substeps", type=int, default=200)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--no-small", action="store_true", help="skip the 64^3 and 128^3 scenes")
    ap.add_argument("--no-copy", action="store_true", help="skip the device-to-device copy measurement")
    ap.add_argument("--serve-rate", type=float, nargs="+", default=[100.0, 300.0, 1000.0, 3000.0, 5000.0],
                    help="scene-step request rates (req/s)")
    ap.add_argument("--serve-overload", type=float, default=20000.0,
                    help="one more rate above saturation (0: none): its served rate is the saturation rate")
    ap.add_argument("--serve-graph-after", type=int, default=8,
                    help="the serving loop replays captured sort periods once the resident scenes are unchanged "
                         "for this many steps (0: eager)")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the timed region, copy a MIB-MiB device buffer 3 times (a kernel of known traffic "
                         "for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); a multi-GPU run (slabs "
              "along z with a ghost-node exchange per substep) is designed but not built", file=sys.stderr)
        return 2

    from shaderintel_amd import scenes
    from shaderintel_amd.mpm import Context, MPM, load_library, _ptr, _check

    ctx = Context(0)
    log(f"device {ctx.name}, {ctx.cus} CUs; scene {args.scene}")
    t0 = time.perf_counter()
    sim = MPM(ctx, **scenes.options(args.scene))
    n = scenes.populate(sim, args.scene)
    log(f"populated {n} particles in {time.perf_counter() - t0:.1f}s")
    L = load_library()
    g = sim.lattice_gravity()

    def run(k: int) -> None:
        _check(L.mpm_step(sim._h, k, _ptr(g), 0.4, 0.9), "step")

    run(args.warmup)
    sim.sync()
    log(f"warmup {args.warmup} substeps done")
    if args.sweep:
        combos = [dict(g2p_waves=w, p2g_waves=p) for w in (3, 4, 5, 6) for p in (4, 5, 6)]
        for c in combos:
            sim.set_tuning(**c)
            run(8)
            sim.sync()
            t0 = time.perf_counter()
            run(args.steps)
            sim.sync()
            dt = time.perf_counter() - t0
            print(json.dumps({"knobs": c, "ms_per_substep": round(dt * 1e3 / args.steps, 4),
                              "particle_substeps_per_s": round(n * args.steps / dt, 1)}), flush=True)
        return 0
    # The timed region. G2P (the dominant kernel) is timed by its own dispatches: the launch records
    # its start and end, so no packet is added to the stream. Every --time-every-th launch is timed
    # (7, prime to the sort period of 4, samples every position in it); a timed launch starts and ends
    # a few microseconds later than an untimed one, so the others run as unprofiled.
    sim.set_profiling(2, every=args.time_every)
    sim.sync()
    t0 = time.perf_counter()
    run(args.steps)
    sim.sync()
    dt = time.perf_counter() - t0
    ktimes = sim.kernel_times()  # {"g2p": (ms, launches), "g2p_p2g": (ms, launches)}
    g2p_ms = ktimes["g2p"][0] + ktimes["g2p_p2g"][0]
    sim.set_profiling(False)
    # Phase breakdown: the next args.steps substeps with events between the phases (those events
    # add packets to the stream, so this pass is reported beside the throughput, not as it).
    sim.set_profiling(True)
    t1 = time.perf_counter()
    run(args.steps)
    sim.sync()
    dt_phases = time.perf_counter() - t1
    phases = sim.phase_times()
    sim.set_profiling(False)
    from shaderintel_amd.mpm import diagnostics
    diag = diagnostics(sim._h)
    log(f"diagnostics after the timed run: {diag}")
    raw = sim.read_particles()
    finite = bool(np.isfinite(raw).all())
    del raw
    log(f"timed {args.steps} substeps in {dt * 1e3:.1f} ms (G2P {g2p_ms:.1f} ms); phase pass "
        f"{dt_phases * 1e3:.1f} ms, phases(ms total) {phases}; finite={finite}")

    value = n * args.steps / dt
    per_launch = {k: v / args.steps for k, v in phases.items()}
    sorts = sum(1 for s in range(args.warmup + args.steps, args.warmup + 2 * args.steps) if s % 4 == 0)
    per_launch["mpm sort"] = phases["mpm sort"] / max(sorts, 1)
    phase_breakdown = {k.replace("mpm ", ""): round(v, 4) for k, v in per_launch.items()}
    # average launch durations of the timed region itself (dispatch-recorded events)
    launch_ms = {"mpm " + k: v[0] / v[1] for k, v in ktimes.items() if v[1]}
    launch_ms.setdefault("mpm p2g", per_launch["mpm p2g"])
    # the dominant kernel: the fused G2P + P2G when the substeps were fused, else G2P
    dom = "mpm g2p_p2g" if "mpm g2p_p2g" in launch_ms else "mpm g2p"
    per_launch[dom] = launch_ms[dom]
    achieved = ALGO_BYTES[dom] * n / (launch_ms[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom.replace(" ", "_"),
                "algo_bytes_per_launch": ALGO_BYTES[dom] * n,
                "launches": {k: {"ms_total": round(v[0], 3), "launches": v[1]} for k, v in ktimes.items()},
                "avg_launch_ms": round(per_launch[dom], 4), "timed_every": args.time_every}
    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3 counter passes
    # (FETCH_SIZE / WRITE_SIZE scaled by a copy kernel of known size) on the same workload
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.scene}.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("kernel", "").split("<")[0] == "k_" + dom.split()[-1]:
            roofline["traffic"] = round(t["traffic_bytes"])
            roofline["traffic_source"] = t["source"]
            # the HBM rate the measured traffic implies at this run's launch time
            traffic_gbs = t["traffic_bytes"] / (launch_ms[dom] * 1e-3) / 1e9
            roofline["traffic_gbs"] = round(traffic_gbs, 1)
            roofline["traffic_frac"] = round(traffic_gbs / HBM_PEAK_GBS, 4)
            if "limiter" in t:
                roofline["limiter"] = t["limiter"]
    substep_bytes = 384 * n
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded jittered blocks (xorshift32, 8 per cell), Gaussian velocity noise",
        "config": {"workload": f"{args.scene}: {sim.dims[0]}^3 grid, {n} particles, "
                               + ", ".join(str(b[0]) if isinstance(b[0], str) else "sand(30deg)"
                                           for b in scenes.SCENES[args.scene]["blocks"]) + ", fp32",
                   "grid": sim.dims, "particles": n, "sort_every": 4, "substeps_timed": args.steps},
        "roofline": roofline,
        "substep_hbm_fraction": round(substep_bytes / (dt / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "phase_ms_per_substep": phase_breakdown,
        "phase_pass_ms_per_substep": round(dt_phases * 1e3 / args.steps, 4),
        "finite": finite,
    }
    sim.destroy()
    if not args.no_copy:
        copy_gbs = copy_bandwidth()
        roofline["copy_gbs"] = round(copy_gbs, 1)
        roofline["frac_of_copy"] = round(achieved / copy_gbs, 4)
        log(f"device-to-device copy: {copy_gbs:.0f} GB/s")
    if not args.no_small:
        out["small_configs"] = gpu_small_configs(ctx)
    if not args.no_batched:
        out["batched"] = batched(ctx, args.batched_substeps)
    if not args.no_serve:
        ga = args.serve_graph_after
        out["serve"] = [serve(ctx, r, args.serve_seconds, args.serve_budget_gb, graph_after=ga) for r in args.serve_rate]
        if args.serve_overload > 0:
            over = serve(ctx, args.serve_overload, args.serve_seconds, args.serve_budget_gb, graph_after=ga)
            over["overload"] = True
            out["serve"].append(over)
            out["serve_saturation_req_s"] = over["scene_steps_per_s"]
    if args.copy_calibration:
        copy_calibration(args.copy_calibration << 20)
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(args.scene, args.cpu_substeps, small=not args.no_small)
        log(f"cpu baseline {out['cpu_baseline']['value']:.3e} particle-substeps/s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
