"""The JavaScript host (shaderintel_amd/js) over the N-API addon.

CPU: the addon loads in Node and the JS host's host-side math (material rows, shape sampling in the
library and in JavaScript for sdf shapes, record packing) produces the oracle's bits; creating a
context without a device raises SIError('no-device').
GPU: a four-material scene driven entirely from JavaScript matches the oracle bit for bit, the
checkMPM law holds, and destroy() is idempotent.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_mpm as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "shaderintel_amd", "js", "mpm_napi.node")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the built addon is missing")


def _u32(a) -> np.ndarray:
    return np.asarray(a, np.uint32)


@needs_node
def test_js_host_helpers_match_oracle(tmp_path):
    out = tmp_path / "helpers.json"
    mode = "nodevice" if not os.path.exists("/dev/kfd") else "device"
    subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "host_helpers.js"), str(out), mode], check=True, timeout=120)
    res = json.loads(out.read_text())
    from shaderintel_amd.mpm import MATERIALS
    for name, m in MATERIALS.items():
        want = O.material_row(m, 2).view(np.uint32)
        assert np.array_equal(_u32(res["materials"][name]), want), name
    box = O.sample_shape(0, [0.1, 0.2, 0.3, 0.4, 0.35, 0.5], 1 / 128, 0.6, 7)
    sph = O.sample_shape(1, [0.5, 0.5, 0.5, 0.15], 1 / 96, 1.0, 3)
    cyl = O.sample_shape(2, [0.5, 0.4, 0.5, 0.1, 0.2], 1 / 96, 0.6, 11)
    assert np.array_equal(_u32(res["samples"]["box"]), box.view(np.uint32).ravel())
    assert np.array_equal(_u32(res["samples"]["sphere"]), sph.view(np.uint32).ravel())
    assert np.array_equal(_u32(res["samples"]["cylinder"]), cyl.view(np.uint32).ravel())
    # the sdf path is JavaScript (xorshift32 in doubles); the sphere sdf over the sphere's bounds must
    # reproduce the library's sphere sampler up to the inside test (|p-c| vs squared distance)
    sdf = _u32(res["samples"]["sdf"]).view(np.float32).reshape(-1, 3)
    assert abs(len(sdf) - len(sph)) <= 3 and len(sdf) > 100
    rec = O.pack_particles(box, [0, 0, 0], 1 / 64, [64, 64, 64], 3.0, 2, [0.1, -0.2, 0.05], (1 / 480) / (1 / 64), 0.01, 12345)
    assert np.array_equal(_u32(res["packed"]), rec.view(np.uint32).ravel())
    if mode == "nodevice":
        assert res["nodevice"] == "no-device|true"


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("frames", [1, 3])
def test_js_scene_matches_oracle(tmp_path, frames):
    print(f"[gpu] JS host scene, {frames} frames", flush=True)
    subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "run_scene.js"), str(tmp_path), str(frames)], check=True,
                   timeout=120)
    summary = json.loads((tmp_path / "summary.json").read_text())
    n = summary["count"]
    before = np.fromfile(tmp_path / "before.f32", np.float32).reshape(n, 32)
    after = np.fromfile(tmp_path / "after.f32", np.float32).reshape(n, 32)
    mats = np.fromfile(tmp_path / "materials.f32", np.float32)
    sc = O.Scene(summary["dims"], n, 2, 4)
    sc.set_materials(mats)
    sc.write(before)
    sc.set_count(n)
    sc.step(8 * frames, np.asarray(summary["gravity"], np.float32))
    want = sc.read(n)
    assert np.array_equal(after.view(np.uint32), want.view(np.uint32))
    pos = np.fromfile(tmp_path / "positions.f32", np.float32).reshape(n, 3)
    assert np.array_equal(pos.view(np.uint32), (after[:, :3].astype(np.float64) * summary["dx"]).astype(np.float32).view(np.uint32))
    after2 = np.fromfile(tmp_path / "after2.f32", np.float32).reshape(n, 32)
    sc.step(1, np.asarray(summary["gravity"], np.float32))
    assert np.array_equal(after2.view(np.uint32), sc.read(n).view(np.uint32))
    assert summary["stepCount"] == 8 * frames + 1
    assert summary["law"]["ok"], summary["law"]
    assert summary["afterDestroy"] == "destroyed"


@needs_node
@pytest.mark.gpu
def test_js_reads_resolve_later(tmp_path):
    """P.read / grid.read / readPoints are Promises settled off the JS thread (napi_async_work):
    a read issued between two steps holds the state after the first (with sortEvery set to 2 first,
    checked against the oracle); P2G into a second grid array with the global path equals the tiled
    P2G into the simulation's grid; and with about a thousand substeps queued ahead of a read, a
    setImmediate callback queued after it runs first."""
    print("[gpu] JS host: reads that resolve later", flush=True)
    subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "async_read.js"), str(tmp_path)], check=True, timeout=120)
    s = json.loads((tmp_path / "summary.json").read_text())
    n = s["count"]
    rec = {k: np.fromfile(tmp_path / f"{k}.f32", np.float32).reshape(n, 32) for k in ("before", "mid", "after")}
    sc = O.Scene(s["dims"], n, 2, 4)
    sc.set_materials(np.fromfile(tmp_path / "materials.f32", np.float32))
    sc.write(rec["before"])
    sc.set_count(n)
    sc.set_sort_every(2)
    g = np.asarray(s["gravity"], np.float32)
    sc.step(8, g)
    assert np.array_equal(rec["mid"].view(np.uint32), sc.read(n).view(np.uint32))
    sc.step(8, g)
    assert np.array_equal(rec["after"].view(np.uint32), sc.read(n).view(np.uint32))
    assert s["sortEvery"] == 2
    assert s["p2gDiffer"] == 0 and s["p2gNonzero"] > 1000
    assert s["badGrid"] == "invalid-argument"
    assert s["order"] == ["timer", "read"], s
    assert s["bigFinite"] == s["bigCount"] > 500000
