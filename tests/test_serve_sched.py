"""The serving loop's host-side scheduler on the CPU with a forced small budget: admission order,
no overtaking, eviction at the end of a lifetime, rejection, latency accounting
(tests/sched_check.cpp against shaderintel_amd/csrc/serve_sched.hpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_scheduler_with_small_budget(tmp_path):
    exe = str(tmp_path / "sched_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'shaderintel_amd', 'csrc')}",
                    os.path.join(ROOT, "tests", "sched_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
