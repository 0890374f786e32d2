"""The grid between the grid update and G2P (the reference's coupling call site, mpm.js:596;
coupling.js:203-221): a host edit of the grid through mpm_substep_begin / mpm_write_grid /
mpm_substep_end gives exactly what the oracle gives for the same edit, and a native hook is called
once per substep with the simulation's grid geometry."""
import numpy as np
import pytest

import oracle_mpm as O
from helpers import bits
from shaderintel_amd.mpm import MPM, GRID_HOOK, SIError, load_library, _ptr

pytestmark = pytest.mark.gpu


def _sim(ctx):
    sim = MPM(ctx, grid=32, maxParticles=40000, substeps=8)
    sim.add({"box": [[0.2, 0.15, 0.2], [0.6, 0.55, 0.6]]}, material="water", velocity=[0.5, -0.3, 0.2], jitter=1.0)
    sim.add({"box": [[0.62, 0.2, 0.3], [0.8, 0.4, 0.5]]}, material="jelly", velocity=[-0.4, 0.0, 0.1], jitter=1.0)
    return sim


def _edit(words, dims):
    """A coupling-like edit: nodes with x < 12 lose half their velocity; mass is kept."""
    v = words.view(np.float32).reshape(dims[2], dims[1], dims[0], 4).copy()
    v[:, :, :12, :3] *= np.float32(0.5)
    return v.reshape(-1, 4).view(np.int32)


def test_split_substep_with_grid_edit_matches_oracle(ctx):
    print("[gpu] split substep + grid edit vs oracle", flush=True)
    sim = _sim(ctx)
    rec = sim.read_particles()
    n = len(rec)
    sc = O.Scene(sim.dims, n, 2, 4)
    sc.set_materials(sim.material_table())
    sc.write(rec)
    sc.set_count(n)
    g = sim.lattice_gravity()
    for k in range(6):
        sim.substep_begin()
        with pytest.raises(SIError):
            sim.step()  # not while a substep is open
        words = sim.read_grid()
        if sc.step_count % 4 == 0:
            sc.sort()
        sc.p2g()
        sc.grid_update(g, 0.4)
        want = sc.grid()
        assert np.array_equal(words, want), f"grid after the update, substep {k}"
        edited = _edit(words, sim.dims)
        sim.write_grid(edited)
        sim.substep_end()
        sc.write_grid(edited)
        sc.g2p(0.9)
        sc.step_count = sc.step_count + 1
        got, exp = sim.read_particles(), sc.read(n)
        assert np.array_equal(bits(got), bits(exp)), f"particles after substep {k}"
    with pytest.raises(SIError):
        sim.substep_end()  # nothing open
    sim.destroy()


def test_native_hook_called_per_substep(ctx):
    print("[gpu] native after-grid-update hook", flush=True)
    sim = _sim(ctx)
    seen = []

    @GRID_HOOK
    def hook(user, view):
        v = view.contents
        seen.append((v.substep, v.step_count, tuple(v.dims), tuple(v.blocks), bool(v.velocity), bool(v.stream)))

    ref = _sim(ctx)
    sim.set_grid_hook(hook)
    sim.step()  # 8 substeps with the hook (no graph capture)
    ref.step()
    assert [s[0] for s in seen] == list(range(8))
    assert [s[1] for s in seen] == list(range(8))
    assert all(s[2] == (32, 32, 32) and s[3] == (8, 8, 8) and s[4] and s[5] for s in seen)
    # a hook that enqueues nothing leaves the results unchanged
    assert np.array_equal(bits(sim.read_particles()), bits(ref.read_particles()))
    sim.set_grid_hook(None)
    sim.step()
    assert len(seen) == 8
    sim.destroy()
    ref.destroy()
