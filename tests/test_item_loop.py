"""The item-persistent kernels' work loop on the CPU (tests/item_loop_check.cpp over
shaderintel_amd/csrc/item_sched.hpp, the functions the kernels call) with the item list of the
8,388,608-particle four-material scene on its 256^3 grid.

The positions are the scene's binned ones moved by up to 2.7 cells per axis (three substeps at the
0.9-cell speed limit: the farthest a particle gets before the next sort), and 1% of them by up to 40
cells (they leave their item's window and take the global-atomic path). Checked: every item is taken
exactly once at six grid sizes in random interleavings; every in-window stencil's LDS tile index,
every tile box (<= 12^3 nodes, 8-bit extents) and every arena offset stays in bounds, and the LDS
indices and boxes of the narrower drift-2 window (10^3 nodes) too. The tile storage order (node
block by node block) is a bijection of every box's nodes, and each node block's chunk, as the grid
update reads it, holds exactly that block's nodes.

The replay of an item slot overwritten before all of a workgroup's waves have read it (a race a
kernel that fetches its next item early can have) shows what such a race does on this item list:
tile boxes spanning two windows, and with the arena check that does not bound the box by the window
tile, arena ends past the arena, i.e. stores outside it."""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np
import pytest

import oracle_mpm as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is missing")
def test_item_loop_bounds_full_scene(tmp_path):
    from shaderintel_amd import scenes
    exe = str(tmp_path / "item_loop_check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-x", "hip",
                    "-I", os.path.join(ROOT, "shaderintel_amd", "csrc"), os.path.join(ROOT, "tests", "item_loop_check.cpp"),
                    "-o", exe], check=True, capture_output=True)
    dims, tbl, rec, g = scenes.host_build("mixed256")
    n = len(rec)
    sc = O.Scene(dims, n, 2, 4)
    sc.set_materials(tbl)
    sc.write(rec)
    sc.set_count(n)
    keys = sc.keys(n)
    del sc
    order = np.argsort(keys >> 6, kind="stable")
    blocks = (keys[order] >> 6).astype(np.uint32)
    heads = np.flatnonzero(np.r_[True, blocks[1:] != blocks[:-1]])
    start = np.r_[heads, n].astype(np.uint32)
    x = rec[order, :3].astype(np.float32)
    del rec
    rng = np.random.default_rng(7)
    x += rng.uniform(-2.7, 2.7, size=x.shape).astype(np.float32)
    far = rng.random(n) < 0.01
    x[far] += rng.uniform(-40, 40, size=(int(far.sum()), 3)).astype(np.float32)
    path = tmp_path / "items.bin"
    with open(path, "wb") as f:
        np.array([n, len(heads)], np.uint32).tofile(f)
        np.array(dims, np.int32).tofile(f)
        np.array([24], np.uint32).tofile(f)
        start.tofile(f)
        blocks[heads].tofile(f)
        np.ascontiguousarray(x).tofile(f)
    out = subprocess.run([exe, str(path)], check=True, capture_output=True, text=True, timeout=600).stdout
    print(out)
    assert out.startswith("ok")
    m = re.search(r"(\d+) item pairs, (\d+) boxes over \d+ nodes, (\d+) with an extent over 255, (\d+) arena ends", out)
    pairs, over, ext, oob = (int(v) for v in m.groups())
    assert pairs > 0 and over > 0 and oob > 0  # the race, where it to happen, writes outside the arena
