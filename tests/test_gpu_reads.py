"""Reads that resolve later and the sortEvery setter, from Python through the C ABI.

mpm_read_*_begin takes its snapshot in stream order when it is called: steps enqueued after it do
not change what mpm_read_finish returns. MPM.sortEvery set between steps changes the sort period from
the next substep on (mpm.js:591: sort when stepCount % sortEvery == 0), bit for bit with the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_mpm as O
from helpers import bits
from shaderintel_amd.mpm import MPM, SIError, load_library, _ptr

pytestmark = pytest.mark.gpu


def _scene(ctx, sort_every=4):
    sim = MPM(ctx, grid=32, maxParticles=40000, substeps=8, sortEvery=sort_every)
    sim.add({"box": [[0.2, 0.15, 0.2], [0.5, 0.55, 0.45]]}, {"material": "water"})
    sim.add({"sphere": [[0.7, 0.4, 0.5], 0.12]}, {"material": "snow", "velocity": [-0.3, 0, 0]})
    rec = sim.read_particles()
    sc = O.Scene(sim.dims, len(rec), sim.ppcAxis, sort_every)
    sc.set_materials(sim.material_table())
    sc.write(rec)
    sc.set_count(len(rec))
    return sim, sc


def test_async_reads_snapshot_in_stream_order(ctx):
    print("[gpu] reads: snapshots in stream order", flush=True)
    sim, sc = _scene(ctx)
    g = sim.lattice_gravity()
    L = load_library()
    assert L.mpm_step(sim._h, 8, _ptr(g), 0.4, 0.9) == 0
    pr, pg = sim.read_particles_async(), sim.read_grid_async()
    later = sim.read_particles_async(100, offset=50)
    assert L.mpm_step(sim._h, 8, _ptr(g), 0.4, 0.9) == 0  # enqueued after the snapshots
    sc.step(8, g)
    assert np.array_equal(bits(pr.result()), bits(sc.read(sim.count)))
    assert np.array_equal(pg.result(), sc.grid())
    assert np.array_equal(bits(later.result()), bits(sc.read(sim.count)[50:150]))
    assert pr.done() and pr.result() is pr.result()
    sc.step(8, g)
    assert np.array_equal(bits(sim.read_particles_async().result()), bits(sc.read(sim.count)))
    # an unfinished read is released with its object
    dropped = sim.read_particles_async()
    del dropped
    with pytest.raises(SIError):
        sim.read_particles_async(10, offset=2**31)
    sim.destroy()


@pytest.mark.parametrize("first,then", [(4, 3), (2, 5)])
def test_sort_every_setter(ctx, first, then):
    print(f"[gpu] reads: sortEvery {first} -> {then}", flush=True)
    sim, sc = _scene(ctx, first)
    g = sim.lattice_gravity()
    L = load_library()
    assert L.mpm_step(sim._h, 5, _ptr(g), 0.4, 0.9) == 0
    sc.step(5, g)
    sim.sortEvery = then
    sc.set_sort_every(then)
    assert sim.sortEvery == then
    for k in (7, 4):
        assert L.mpm_step(sim._h, k, _ptr(g), 0.4, 0.9) == 0
        sc.step(k, g)
    got, want = sim.read_particles(), sc.read(sim.count)
    assert np.array_equal(bits(got), bits(want))
    with pytest.raises(SIError):
        sim.sortEvery = 0
    sim.destroy()
