"""The C ABI: every function include/mpm.h declares is exported by libmpm_hip.so, and the library
loads on a host without a GPU (no compute calls here). The host-only helpers run on the CPU."""
import ctypes
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "shaderintel_amd", "lib", "libmpm_hip.so")


def _declared():
    src = open(os.path.join(ROOT, "include", "mpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = _declared()
    for required in ("mpm_ctx_create", "mpm_sim_create", "mpm_set_materials", "mpm_set_colliders",
                     "mpm_write_particles", "mpm_set_count", "mpm_step", "mpm_p2g", "mpm_sort", "mpm_rescale",
                     "mpm_read_particles", "mpm_read_grid", "mpm_sync", "mpm_sim_destroy", "mpm_status_string",
                     "mpm_last_error_message", "mpm_scan_u32", "mpm_counting_sort", "mpm_pool_create",
                     "mpm_pool_step"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(mpm_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    L = ctypes.CDLL(LIB)
    for n in _declared():
        getattr(L, n)


def test_host_helpers_without_gpu():
    from shaderintel_amd.mpm import load_library, SIError
    L = load_library()
    assert L.mpm_status_string(1) == b"invalid-argument"
    assert L.mpm_status_string(5) == b"out-of-memory"
    out = np.zeros(12, np.float32)
    L.mpm_material_row(1, 1.0, 0.35, 0.3, 0.0, 10.0, 0.025, 0.0075, 30.0, 2, out.ctypes.data_as(ctypes.c_void_p))
    assert out[1] == np.float32(1 / 8)
    assert SIError("x", "no-device").code == "no-device"


def test_record_slot_tables_agree():
    """The stream slot of each reference record float is stated three times: the public header
    (MPM_REC_SLOT), the kernels (kRecSlot) and the Python mirror (REC_SLOT). They must agree and be a
    permutation of the 32 record floats."""
    from shaderintel_amd.mpm import REC_SLOT
    hdr = open(os.path.join(ROOT, "include", "mpm.h")).read()
    m = re.search(r"#define MPM_REC_SLOT\s*\\\s*\{([^}]*)\}", hdr)
    assert m
    header = [int(v) for v in m.group(1).split(",")]
    ker = open(os.path.join(ROOT, "shaderintel_amd", "csrc", "mpm_kernels.hpp")).read()
    m = re.search(r"kRecSlot\[32\] = \{(.*?)\};", ker, flags=re.S)
    assert m
    kernels = [int(v) for v in re.findall(r"\b\d+\b", re.sub(r"//[^\n]*", "", m.group(1)))]
    assert header == kernels == list(REC_SLOT)
    assert sorted(header) == list(range(32))
    assert header[:4] == [0, 1, 2, 3]  # position and material stay in stream 0 (the render layers' slots)
