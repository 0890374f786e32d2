"""Writes tests/golden/mpm_small.npz: small input / output vectors of the MLS-MPM substep. The i32
P2G words come from tests/literal.py, the node-by-node numpy restatement of mpm.js:318-349 that
shares no code with the oracle, and are checked equal to the oracle's before anything is written;
the records after 1 and 6 substeps and the grid after the 6th update come from the oracle (the CPU
restatement under oracle/). The CPU suite checks that the oracle still reproduces them; the GPU
suite checks the HIP path against them without the oracle.

Scene: 16^3 grid, ppc 2, bound 3, sort every 4; water, jelly, snow and sand boxes (2,304
particles) with randomized v, C, tau and a sheared F; lattice gravity (0, -2.725e-3, 0).
Contents: rec0 (input records), grid_p2g (i32 words after one P2G of rec0), rec1 / rec6 (records
after 1 and 6 substeps), grid6 (f32 bits of the grid velocities after the 6th grid update),
materials (8 x 12 table), g (lattice gravity).

Run: python tests/golden/make_golden.py   (needs oracle/_build/liboracle_mpm.so: make -C oracle)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle_mpm as O  # noqa: E402

sys.path.insert(0, os.path.join(HERE, ".."))
import literal  # noqa: E402

MATS = [{"kind": 0, "density": 1.0, "stiffness": 1.0, "viscosity": 0.0},
        {"kind": 1, "density": 1.0, "stiffness": 0.35, "poisson": 0.3},
        {"kind": 2, "density": 0.4, "stiffness": 0.4, "poisson": 0.2, "hardening": 10, "compression": 0.025,
         "stretch": 0.0075},
        {"kind": 3, "density": 1.6, "stiffness": 0.6, "poisson": 0.3, "friction": 30}]
DIMS = [16, 16, 16]


def build():
    dx, dt = 1 / 16, 1 / 480
    boxes = [(0, [3, 3, 3], [7, 9, 7]), (1, [9, 4, 3], [13, 8, 7]), (2, [3, 4, 9], [7, 8, 13]), (3, [9, 3, 9], [13, 7, 13])]
    recs = []
    seed = 1
    for mat, lo, hi in boxes:
        pts = O.sample_shape(0, [l * dx for l in lo] + [h * dx for h in hi], dx / 2, 1.0, seed)
        recs.append(O.pack_particles(pts, [0, 0, 0], dx, DIMS, 3.0, mat, [0.3, -0.2, 0.1], dt / dx))
        seed += len(pts)
    rec = np.concatenate(recs)
    rng = np.random.default_rng(2024)
    n = len(rec)
    rec[:, 4:7] += (rng.random((n, 3)) - 0.5) * 0.04
    for j in range(3):
        rec[:, 8 + 4 * j:11 + 4 * j] = (rng.random((n, 3)) - 0.5) * 0.05
    for k in (11, 15, 19, 23, 27, 31):
        rec[:, k] = (rng.random(n) - 0.5) * 0.02
    solid = rec[:, 3].view(np.uint32) != 0
    F = np.eye(3)[None] + (rng.random((n, 3, 3)) - 0.5) * 0.2
    for j in range(3):
        rec[solid, 20 + 4 * j:23 + 4 * j] = F[solid, :, j]
    tbl = np.zeros((8, 12), np.float32)
    for k, m in enumerate(MATS):
        tbl[k] = O.material_row(m, 2)
    g = np.array([0.0, np.float32(-9.81 * dt * dt / (1 / 64)), 0.0], np.float32)
    return rec.astype(np.float32), tbl, g


def run(rec, tbl, g):
    n = len(rec)
    sc = O.Scene(DIMS, n, 2, 4)
    sc.set_materials(tbl)
    sc.write(rec)
    sc.set_count(n)
    sc.p2g()
    grid_p2g = sc.grid()
    sc.step(1, g)
    rec1 = sc.read(n)
    sc.step(5, g)
    return grid_p2g, rec1, sc.read(n), sc.grid()


def main():
    rec, tbl, g = build()
    grid_p2g, rec1, rec6, grid6 = run(rec, tbl, g)
    per_node = literal.p2g(rec, tbl, 0.5, DIMS)
    if not np.array_equal(per_node, grid_p2g):
        raise SystemExit(f"oracle P2G differs from the per-node restatement in {int((per_node != grid_p2g).sum())} words")
    grid_p2g = per_node
    np.savez_compressed(os.path.join(HERE, "mpm_small.npz"), rec0=rec, materials=tbl, g=g, grid_p2g=grid_p2g,
                        rec1=rec1, rec6=rec6, grid6=grid6)
    print(f"wrote mpm_small.npz: {len(rec)} particles")


if __name__ == "__main__":
    main()
