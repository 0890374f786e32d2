"""The benchmark's contract on the CPU: one GPU only (a request for more exits with a message before
anything touches a device), the algorithmic byte counts its roofline credits, and the committed HBM
traffic record it reports beside them."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line for a run that did not happen


def test_algorithmic_bytes():
    # P2G reads the eight 16-byte streams of a record; G2P reads and writes them; the fused kernel is
    # one of each, the whole particle-substep
    assert bench.ALGO_BYTES["mpm p2g"] == 8 * 16
    assert bench.ALGO_BYTES["mpm g2p"] == 2 * 8 * 16
    assert bench.ALGO_BYTES["mpm g2p_p2g"] == bench.ALGO_BYTES["mpm p2g"] + bench.ALGO_BYTES["mpm g2p"]
    assert bench.HBM_PEAK_GBS == 8000.0


def test_percentiles():
    assert bench.pctl([], 50) is None
    assert bench.pctl([1.0, 2.0, 3.0], 50) == 2.0
    assert bench.pctl(list(range(101)), 99) == 99.0


@pytest.mark.parametrize("scene,particles", [("mixed256", 8388608)])
def test_traffic_record(scene, particles):
    with open(os.path.join(ROOT, "profiles", f"traffic_{scene}.json")) as f:
        t = json.load(f)
    assert t["workload"] == scene and t["kernel"] == "k_g2p_p2g" and t["instance"].startswith("k_g2p_p2g<")
    assert t["limiter"]["unit"] == "VALU" and 0 < t["limiter"]["valu_busy"] <= 1
    assert t["algo_bytes"] == bench.ALGO_BYTES["mpm g2p_p2g"] * particles
    assert t["traffic_bytes"] == pytest.approx(t["read_bytes"] + t["write_bytes"])
    # the fused kernel writes each record once (128 B per particle) and its tiles; its reads are
    # what the caches do not catch, so the measured traffic stays under the algorithmic bytes
    assert 128 * particles <= t["write_bytes"] and t["traffic_bytes"] < t["algo_bytes"]
    # counter scales from copies of known size: FETCH_SIZE counts half the bytes on gfx950
    assert t["fetch_scale"] == pytest.approx(2.0, rel=0.01) and t["write_scale"] == pytest.approx(1.0, rel=0.01)
    assert os.path.exists(os.path.join(ROOT, t["source"]))
