"""Build the native parts of shaderintel_amd in-tree.

- ``lib/libmpm_hip.so``: the HIP library behind ``include/mpm.h`` (kernels for gfx950 only).
- ``js/mpm_napi.node``: the Node N-API addon that exposes the C ABI to the JavaScript host.

Run ``python -m shaderintel_amd.build`` (or ``__graft_entry__.build()``). Cross-compiles without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libmpm_hip.so")
ADDON = os.path.join(PKG, "js", "mpm_napi.node")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: every float operation rounds as written (the kernels are bit-reproducible).
# -packed-fp32-ops: packed fp32 instructions disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC",
    "-Wall",
]
SOURCES = ["prims.hip", "mpm_kernels.hip", "mpm_runtime.cpp", "host_tools.cpp", "serve.cpp"]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"build failed ({res.returncode}): {' '.join(cmd)}\n{out}")


def build_library() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    obj_dir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(obj_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") or src == "mpm_runtime.cpp" else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_checks() -> str:
    """The GPU math check library the tests call (tests/gpu_math_check.hip)."""
    out_dir = os.path.join(ROOT, "tests", "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libgpu_math_check.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{CSRC}", os.path.join(ROOT, "tests", "gpu_math_check.hip"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_addon() -> str | None:
    """Build the N-API addon if node headers are present (returns None otherwise)."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = ADDON + ".tmp"
    _run([cxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", f"-I{inc}", f"-I{os.path.join(ROOT, 'include')}",
          "-DNAPI_VERSION=8", os.path.join(PKG, "js", "addon.cc"), "-o", tmp,
          f"-L{LIB_DIR}", "-lmpm_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, ADDON)
    return ADDON


def build_oracle() -> str | None:
    """Compile the CPU oracle (test infrastructure) with its own Makefile."""
    odir = os.path.join(ROOT, "oracle")
    if not os.path.exists(os.path.join(odir, "Makefile")):
        return None
    _run(["make", "-s", "-C", odir])
    return os.path.join(odir, "_build", "liboracle_mpm.so")


def main() -> None:
    print("lib:", build_library())
    print("addon:", build_addon())
    print("oracle:", build_oracle())


if __name__ == "__main__":
    sys.exit(main())
