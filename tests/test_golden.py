"""Golden vectors (tests/golden/mpm_small.npz, written by tests/golden/make_golden.py): the oracle
still reproduces them (CPU), and the HIP path reproduces them bit for bit without the oracle (GPU):
one P2G, then 1 and 6 substeps of a four-material scene with randomized transfer terms."""
import os

import numpy as np
import pytest

from helpers import bits

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "mpm_small.npz"))


def test_oracle_reproduces_golden():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden
    grid_p2g, rec1, rec6, grid6 = make_golden.run(G["rec0"], G["materials"], G["g"])
    rec0, _, _ = make_golden.build()
    assert np.array_equal(bits(rec0), bits(G["rec0"]))
    assert np.array_equal(grid_p2g, G["grid_p2g"])
    assert np.array_equal(bits(rec1), bits(G["rec1"]))
    assert np.array_equal(bits(rec6), bits(G["rec6"]))
    assert np.array_equal(grid6, G["grid6"])


@pytest.mark.gpu
def test_hip_reproduces_golden(ctx):
    print("[gpu] golden vectors mpm_small.npz", flush=True)
    from shaderintel_amd.mpm import MPM, load_library, _ptr
    rec0 = G["rec0"]
    n = len(rec0)
    sim = MPM(ctx, grid=16, maxParticles=n, substeps=8)
    L = load_library()
    assert L.mpm_set_materials(sim._h, _ptr(np.ascontiguousarray(G["materials"])), 8) == 0
    sim.write_particles(rec0)
    assert L.mpm_set_count(sim._h, n) == 0
    sim.p2g()
    assert np.array_equal(sim.read_grid(), G["grid_p2g"])
    g = np.ascontiguousarray(G["g"])
    assert L.mpm_step(sim._h, 1, _ptr(g), 0.4, 0.9) == 0
    assert np.array_equal(bits(sim.read_particles()), bits(G["rec1"]))
    assert L.mpm_step(sim._h, 5, _ptr(g), 0.4, 0.9) == 0
    assert np.array_equal(bits(sim.read_particles()), bits(G["rec6"]))
    assert np.array_equal(sim.read_grid(), G["grid6"])
    sim.destroy()
