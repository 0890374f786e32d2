"""Python mirror of ShaderIntel's ``MPM`` world API over the C ABI of ``libmpm_hip.so``.

It keeps the reference's names, argument meanings and error behaviour (src/physics/mpm.js:82-680,
types/shaderintel.d.ts:813-900): ``MPM(ctx, maxParticles=…, grid=…, size=…, …)``, ``add``,
``clear``, ``setColliders``, ``step``, ``p2g``, ``_sort``, ``readPositions`` and the fields
``count, dims, dx, origin, bound, numNodes, materials, ppcAxis, substeps, timeScale, sortEvery,
gravity, substepDt, stepCount``. Errors are raised as :class:`SIError` with the reference's
``code`` strings. Every method runs on the GPU through the HIP library; there is no CPU path and
a missing library raises at import of the first context.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
from typing import Iterable, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libmpm_hip.so")

FLUID, JELLY, SNOW, SAND = 0, 1, 2, 3
MPM_KINDS = {"fluid": FLUID, "jelly": JELLY, "elastic": JELLY, "snow": SNOW, "sand": SAND}
MAX_DENSITY = 50  # mpm.js:34
# Device particle streams (MPM_REC_SLOT in include/mpm.h): float f of a reference record (mpm.js:54-63)
# is component REC_SLOT[f] % 4 of stream REC_SLOT[f] // 4.
REC_SLOT = (0, 1, 2, 3, 4, 5, 6, 22, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 24, 25, 26, 7, 27, 28, 29, 20, 30,
            31, 23, 21)

# mpm.js:38-46
MATERIALS = {
    "water": {"kind": FLUID, "density": 1, "stiffness": 1.0, "viscosity": 0.0, "color": [0.12, 0.42, 0.95]},
    "oil": {"kind": FLUID, "density": 0.9, "stiffness": 0.9, "viscosity": 0.3, "color": [0.9, 0.62, 0.12]},
    "honey": {"kind": FLUID, "density": 1.4, "stiffness": 0.9, "viscosity": 1.0, "color": [0.95, 0.6, 0.05]},
    "jelly": {"kind": JELLY, "density": 1, "stiffness": 0.35, "poisson": 0.3, "color": [0.95, 0.3, 0.35]},
    "rubber": {"kind": JELLY, "density": 1.2, "stiffness": 0.7, "poisson": 0.45, "color": [0.25, 0.25, 0.28]},
    "snow": {"kind": SNOW, "density": 0.4, "stiffness": 0.6, "poisson": 0.2, "hardening": 10, "compression": 0.025,
             "stretch": 0.0075, "color": [0.95, 0.96, 1.0]},
    "sand": {"kind": SAND, "density": 1.6, "stiffness": 0.6, "poisson": 0.3, "friction": 33,
             "color": [0.86, 0.68, 0.42]},
}

_CODES = {0: "ok", 1: "invalid-argument", 2: "unsupported", 3: "no-device", 4: "error", 5: "out-of-memory",
          6: "destroyed"}


class SIError(Exception):
    """The reference's error class (src/core/errors.js:5-11): ``code`` is one of 'invalid-argument',
    'unsupported', 'no-device', 'error', plus 'out-of-memory' and 'destroyed'."""

    def __init__(self, msg: str, code: str = "error"):
        super().__init__(msg if msg.startswith("ShaderIntel") else f"ShaderIntel: {msg}")
        self.code = code


class _Config(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 3), ("max_particles", ctypes.c_uint32), ("ppc_axis", ctypes.c_int32),
                ("bound", ctypes.c_float), ("sort_every", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class _PoolConfig(ctypes.Structure):
    _fields_ = [("max_particles", ctypes.c_uint32), ("max_scenes", ctypes.c_uint32),
                ("max_node_blocks", ctypes.c_uint32), ("max_dim", ctypes.c_int32), ("sort_every", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class _SceneDesc(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 3), ("ppc_axis", ctypes.c_int32), ("bound", ctypes.c_float),
                ("gravity", ctypes.c_float * 3), ("num_materials", ctypes.c_uint32), ("materials", ctypes.c_void_p),
                ("records", ctypes.c_void_p), ("count", ctypes.c_uint32)]


class _Diagnostics(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in ("work_items", "s_blocks", "d_blocks", "tile_nodes_used",
                                                "tile_nodes_capacity", "particles")]


def diagnostics(handle, pool: bool = False) -> dict:
    """Device counters of the last binning and P2G of a simulation (or pool) handle."""
    d = _Diagnostics()
    L = load_library()
    _check((L.mpm_pool_diagnostics if pool else L.mpm_diagnostics_of)(handle, ctypes.byref(d)), "diagnostics")
    return {k: getattr(d, k) for k, _ in _Diagnostics._fields_}


class _GridLaw(ctypes.Structure):
    _fields_ = [("mass_words", ctypes.c_int64), ("p2g", ctypes.c_int64 * 3), ("momentum", ctypes.c_int64 * 3),
                ("gravity", ctypes.c_int64 * 3), ("boundary", ctypes.c_int64 * 3), ("nodes", ctypes.c_uint64),
                ("updates", ctypes.c_uint64)]


def grid_law(handle, pool: bool = False, reset: bool = False) -> dict:
    """The grid update's momentum bookkeeping of a simulation (or pool) handle (mpm_grid_law): the raw
    64-bit sums ("raw", 14 ints: mass words, p2g[3], momentum[3], gravity[3], boundary[3], nodes) and
    the same in lattice units (mass, p2g, momentum, gravity, boundary as float64)."""
    d = _GridLaw()
    L = load_library()
    _check((L.mpm_pool_read_grid_law if pool else L.mpm_read_grid_law)(handle, ctypes.byref(d), 1 if reset else 0),
           "grid_law")
    raw = [d.mass_words, *d.p2g, *d.momentum, *d.gravity, *d.boundary, d.nodes]
    return {"raw": [int(x) for x in raw], "updates": int(d.updates), "nodes": int(d.nodes),
            "mass": d.mass_words / 2.0 ** 22, "p2g": np.array(d.p2g[:], np.float64) / 2.0 ** 22,
            "momentum": np.array(d.momentum[:], np.float64) / 2.0 ** 32,
            "gravity": np.array(d.gravity[:], np.float64) / 2.0 ** 32,
            "boundary": np.array(d.boundary[:], np.float64) / 2.0 ** 32}


class GridView(ctypes.Structure):
    """mpm_grid_view (include/mpm.h): what an after-grid-update hook receives."""
    _fields_ = [("velocity", ctypes.c_void_p), ("block_base", ctypes.c_uint32), ("blocks", ctypes.c_int32 * 3),
                ("dims", ctypes.c_int32 * 3), ("s_blocks", ctypes.c_void_p), ("d_blocks", ctypes.c_void_p),
                ("counts", ctypes.c_void_p), ("stream", ctypes.c_void_p), ("substep", ctypes.c_uint32),
                ("step_count", ctypes.c_uint64)]


GRID_HOOK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(GridView))


class _ServerConfig(ctypes.Structure):
    _fields_ = [("substeps", ctypes.c_uint32), ("floor_friction", ctypes.c_float), ("vmax", ctypes.c_float),
                ("graph_after", ctypes.c_uint32)]


class _ServerCounters(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("waiting", ctypes.c_uint64), ("running", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("particle_substeps", ctypes.c_uint64), ("resident_particles", ctypes.c_uint32),
                ("error", ctypes.c_int32), ("scene_steps", ctypes.c_uint64)]


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libmpm_hip.so and declare its signatures (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SIError(f"libmpm_hip.so is not built ({path}); run `python -m shaderintel_amd.build`", "unsupported")
    L = ctypes.CDLL(path)
    vp, u32, i32, u64, f32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "mpm_ctx_create": (i32, [ctypes.c_int, P(vp)]),
        "mpm_ctx_destroy": (None, [vp]),
        "mpm_ctx_info": (i32, [vp, ctypes.c_char_p, ctypes.c_size_t, P(i32)]),
        "mpm_sim_create": (i32, [vp, P(_Config), P(vp)]),
        "mpm_sim_destroy": (None, [vp]),
        "mpm_set_materials": (i32, [vp, vp, u32]),
        "mpm_set_colliders": (i32, [vp, vp, u32]),
        "mpm_write_particles": (i32, [vp, vp, u32, u32]),
        "mpm_set_count": (i32, [vp, u32]),
        "mpm_get_count": (i32, [vp, P(u32)]),
        "mpm_set_step_count": (i32, [vp, u64]),
        "mpm_get_step_count": (i32, [vp, P(u64)]),
        "mpm_step": (i32, [vp, u32, vp, f32, f32]),
        "mpm_p2g": (i32, [vp, i32]),
        "mpm_p2g_to": (i32, [vp, i32, vp]),
        "mpm_sort": (i32, [vp]),
        "mpm_rescale": (i32, [vp, f32]),
        "mpm_read_particles": (i32, [vp, vp, u32, u32]),
        "mpm_read_grid": (i32, [vp, vp]),
        "mpm_sync": (i32, [vp]),
        "mpm_set_profiling": (i32, [vp, i32]),
        "mpm_set_profiling_every": (i32, [vp, u32]),
        "mpm_kernel_times": (i32, [vp, vp, vp]),
        "mpm_set_tuning": (i32, [vp, i32, i32]),
        "mpm_phase_times": (i32, [vp, vp]),
        "mpm_scan_u32": (i32, [vp, vp, vp, u64]),
        "mpm_counting_sort": (i32, [vp, vp, u64, u32, vp, vp]),
        "mpm_material_row": (None, [i32, f64, f64, f64, f64, f64, f64, f64, f64, i32, vp]),
        "mpm_sample_shape": (ctypes.c_int64, [i32, vp, f64, f64, u32, vp, ctypes.c_int64]),
        "mpm_pack_particles": (None, [vp, ctypes.c_int64, vp, f64, vp, f64, u32, vp, f64, f64, u32, vp]),
        "mpm_pool_bytes": (ctypes.c_size_t, [P(_PoolConfig)]),
        "mpm_pool_create": (i32, [vp, P(_PoolConfig), P(vp)]),
        "mpm_pool_destroy": (None, [vp]),
        "mpm_pool_add_scene": (i32, [vp, P(_SceneDesc), P(u32)]),
        "mpm_pool_remove_scene": (i32, [vp, u32]),
        "mpm_pool_step": (i32, [vp, u32, f32, f32]),
        "mpm_pool_scene_count": (i32, [vp, u32, P(u32)]),
        "mpm_pool_read_scene": (i32, [vp, u32, vp, u32, P(u32)]),
        "mpm_pool_info": (i32, [vp, P(u32), P(u32), P(u32)]),
        "mpm_pool_sync": (i32, [vp]),
        "mpm_pool_use_graphs": (i32, [vp, i32]),
        "mpm_pool_set_tuning": (i32, [vp, i32, i32]),
        "mpm_pool_autotune": (i32, [vp, u32, P(i32), P(i32)]),
        "mpm_pool_tuning": (i32, [vp, i32, P(i32)]),
        "mpm_pool_diagnostics": (i32, [vp, P(_Diagnostics)]),
        "mpm_set_grid_law": (i32, [vp, i32]),
        "mpm_read_grid_law": (i32, [vp, P(_GridLaw), i32]),
        "mpm_pool_set_grid_law": (i32, [vp, i32]),
        "mpm_pool_read_grid_law": (i32, [vp, P(_GridLaw), i32]),
        "mpm_scan_u32_noalloc": (i32, [vp, vp, vp, u64, vp]),
        "mpm_set_grid_hook": (i32, [vp, GRID_HOOK, vp]),
        "mpm_read_points": (i32, [vp, vp, vp, u32, u32, vp, f64, f64]),
        "mpm_read_particles_begin": (i32, [vp, u32, u32, P(vp)]),
        "mpm_read_grid_begin": (i32, [vp, P(vp)]),
        "mpm_read_points_begin": (i32, [vp, u32, u32, vp, f64, f64, i32, P(vp)]),
        "mpm_read_finish": (i32, [vp, vp]),
        "mpm_set_sort_every": (i32, [vp, u32]),
        "mpm_particle_streams": (i32, [vp, vp, P(u32)]),
        "mpm_spatial_grid": (i32, [vp, vp, u32, vp, vp, f64, vp, u32, vp, vp, vp, vp, u64, P(u64)]),
        "mpm_substep_begin": (i32, [vp, vp, f32]),
        "mpm_substep_end": (i32, [vp, f32]),
        "mpm_write_grid": (i32, [vp, vp]),
        "mpm_diagnostics_of": (i32, [vp, P(_Diagnostics)]),
        "mpm_pool_add_template": (i32, [vp, P(_SceneDesc), P(u32)]),
        "mpm_pool_remove_template": (i32, [vp, u32]),
        "mpm_pool_add_scene_from_template": (i32, [vp, u32, P(u32)]),
        "mpm_server_create": (i32, [vp, P(_ServerConfig), P(vp)]),
        "mpm_server_submit": (i32, [vp, u32, u32, P(u64)]),
        "mpm_server_submit_result": (i32, [vp, u32, u32, vp, u32, P(u64)]),
        "mpm_server_wait": (i32, [vp, u64, i32, P(i32), P(u32)]),
        "mpm_server_stats": (i32, [vp, P(_ServerCounters)]),
        "mpm_server_latencies": (i32, [vp, i32, vp, u64, P(u64)]),
        "mpm_server_stop": (i32, [vp, i32]),
        "mpm_server_destroy": (None, [vp]),
        "mpm_status_string": (ctypes.c_char_p, [i32]),
        "mpm_last_error_message": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(st: int, what: str = "") -> None:
    if st != 0:
        L = load_library()
        msg = L.mpm_last_error_message().decode(errors="replace")
        raise SIError(f"{what}: {msg}" if what else msg, _CODES.get(st, "error"))


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    """Pointer to an array's data that keeps the array alive while the pointer object lives (so a
    temporary passed straight into a ctypes call is not freed before the call runs)."""
    return a.ctypes.data_as(ctypes.c_void_p)


def spatial_grid(ctx, points: np.ndarray, lo, hi, cell_size: float, queries: np.ndarray):
    """SpatialGrid build + neighbour ranges (src/prims/grid.js:24-117) on the device: returns
    (order, start, offsets, neighbors) with neighbors[offsets[q]:offsets[q+1]] the sorted-order
    indices in the 27 cells around query q."""
    L = load_library()
    pts = np.zeros((len(points), 4), np.float32)
    pts[:, :3] = np.asarray(points, np.float32).reshape(-1, 3)
    qs = np.zeros((len(queries), 4), np.float32)
    qs[:, :3] = np.asarray(queries, np.float32).reshape(-1, 3)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    dims = [max(1, int(np.ceil((hi[a] - lo[a]) / cell_size))) for a in range(3)]
    order = np.zeros(len(pts), np.uint32)
    start = np.zeros(dims[0] * dims[1] * dims[2] + 1, np.uint32)
    offsets = np.zeros(len(qs) + 1, np.uint32)
    total = ctypes.c_uint64()
    args = [ctx._h, _ptr(pts), len(pts), _ptr(lo), _ptr(hi), float(cell_size), _ptr(qs), len(qs), _ptr(order),
            _ptr(start), _ptr(offsets)]
    _check(L.mpm_spatial_grid(*args, None, 0, ctypes.byref(total)), "spatial_grid")
    nb = np.zeros(max(total.value, 1), np.uint32)
    _check(L.mpm_spatial_grid(*args, _ptr(nb), total.value, ctypes.byref(total)), "spatial_grid")
    return order, start, offsets, nb[:total.value]


class GArray:
    """A typed array of n elements (gpu.array, src/core/context.js): here 'vec4i', the target of
    MPM.p2g(P, forceGlobal, grid); read() returns its int32 words (4 per element)."""

    KINDS = {"vec4i": (np.int32, 4), "i32": (np.int32, 1), "u32": (np.uint32, 1), "f32": (np.float32, 1),
             "vec4f": (np.float32, 4)}

    def __init__(self, kind: str, n: int):
        if kind not in self.KINDS:
            raise SIError(f"gpu.array: unknown type {kind!r}", "invalid-argument")
        dt, w = self.KINDS[kind]
        self.kind, self.length = kind, int(n)
        self._data = np.zeros(self.length * w, dt)

    def read(self) -> np.ndarray:
        return self._data.copy()

    def write(self, data, offset: int = 0) -> None:
        a = np.asarray(data, self._data.dtype).reshape(-1)
        self._data[offset:offset + a.size] = a


class Context:
    """A HIP device (stands in for the reference's WebGPU Context from SI.init, context.js:32-100)."""

    def __init__(self, device: int = 0):
        L = load_library()
        h = ctypes.c_void_p()
        _check(L.mpm_ctx_create(device, ctypes.byref(h)), "init")
        self._h = h
        buf = ctypes.create_string_buffer(256)
        cus = ctypes.c_int32()
        _check(L.mpm_ctx_info(h, buf, 256, ctypes.byref(cus)))
        self.name = buf.value.decode()
        self.cus = cus.value

    def mpm(self, o: dict | None = None, **kw) -> "MPM":  # gpu.mpm(o), src/index.js:87
        return MPM(self, **{**(o or {}), **kw})

    def array(self, kind: str, n: int) -> "GArray":
        """gpu.array(type, n) for the grids MPM.p2g can target: 'vec4i' of n nodes (host-backed here;
        the library writes it, read() returns it)."""
        return GArray(kind, n)

    def scan(self, a: np.ndarray) -> np.ndarray:
        """Scan.run (src/prims/scan.js:292-306): exclusive prefix sum of u32 (wrapping)."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        out = np.empty_like(a)
        _check(load_library().mpm_scan_u32(self._h, _ptr(a), _ptr(out), a.size), "Scan.run")
        return out

    def counting_sort(self, keys: np.ndarray, num_buckets: int) -> tuple[np.ndarray, np.ndarray]:
        """SpatialGrid.build (src/prims/grid.js:70-79): returns (order, start)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        order = np.empty(keys.size, np.uint32)
        start = np.empty(num_buckets + 1, np.uint32)
        _check(load_library().mpm_counting_sort(self._h, _ptr(keys), keys.size, num_buckets, _ptr(order), _ptr(start)),
               "SpatialGrid.build")
        return order, start

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            load_library().mpm_ctx_destroy(self._h)
            self._h = None


def material_row(m: dict, ppc_axis: int) -> np.ndarray:
    """_uploadMaterials row (mpm.js:170-196) computed by the library's host helper."""
    out = np.zeros(12, np.float32)
    load_library().mpm_material_row(
        int(m["kind"]), float(m.get("density", 1) if m.get("density") is not None else 1),
        float(m["stiffness"]) if m.get("stiffness") is not None else 1.0,
        float(m["poisson"]) if m.get("poisson") is not None else 0.3,
        float(m["viscosity"]) if m.get("viscosity") is not None else 0.0,
        float(m["hardening"]) if m.get("hardening") is not None else 10.0,
        float(m["compression"]) if m.get("compression") is not None else 0.025,
        float(m["stretch"]) if m.get("stretch") is not None else 0.0075,
        float(m["friction"]) if m.get("friction") is not None else 33.0, int(ppc_axis), _ptr(out))
    return out


def sample_shape(shape: dict, spacing: float, jitter: float, seed: int) -> np.ndarray:
    """sampleShape (mpm.js:683-714) for box / sphere / cylinder; float32 xyz points (world)."""
    if "box" in shape:
        lo, hi = shape["box"]
        kind, prm = 0, [*lo, *hi]
    elif "sphere" in shape:
        c, r = shape["sphere"]
        kind, prm = 1, [*c, r]
    elif "cylinder" in shape:
        c, r, h = shape["cylinder"]
        kind, prm = 2, [*c, r, h]
    else:
        raise SIError("MPM.add: shape needs box | sphere | cylinder | points", "invalid-argument")
    p = np.asarray(prm + [0.0] * (6 - len(prm)), np.float64)
    L = load_library()
    n = L.mpm_sample_shape(kind, _ptr(p), spacing, jitter, seed & 0xFFFFFFFF, None, 0)
    out = np.empty((max(n, 0), 3), np.float32)
    if n > 0:
        L.mpm_sample_shape(kind, _ptr(p), spacing, jitter, seed & 0xFFFFFFFF, _ptr(out), n)
    return out


class PendingRead:
    """A read enqueued on a simulation's stream (mpm_read_*_begin); result() waits for it, copies it
    into a host array and returns that array. It holds device staging memory until result() runs."""

    def __init__(self, handle: ctypes.c_void_p, out: np.ndarray, what: str):
        self._h, self._out, self._what = handle, out, what

    def done(self) -> bool:
        return self._h is None

    def result(self) -> np.ndarray:
        if self._h is not None:
            h, self._h = self._h, None
            _check(load_library().mpm_read_finish(h, _ptr(self._out)), self._what)
        return self._out

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            try:
                load_library().mpm_read_finish(self._h, None)
            except Exception:
                pass


class MPM:
    """MLS-MPM simulation on one MI355X (src/physics/mpm.js:82-680)."""

    def __init__(self, ctx: Context, maxParticles: int = 262144, grid: int = 64, size=1, origin=None, gravity=None,
                 substeps: int = 8, timeScale: float = 1, sortEvery: int = 4, ppcAxis: int = 2):
        self.ctx = ctx
        self.maxParticles = int(maxParticles or 262144)
        size = list(size) if isinstance(size, (list, tuple)) else [size or 1] * 3
        res = grid or 64
        self.dx = max(size) / res
        self.dims = [max(8, round(s / self.dx)) for s in size]
        self.size = [d * self.dx for d in self.dims]
        self.origin = list(origin or [0, 0, 0])
        self.gravity = list(gravity or [0, -9.81, 0])
        self._substeps = max(1, round(substeps or 8))
        self._timeScale = 1 if timeScale is None else timeScale
        self._sortEvery = sortEvery or 4
        self.ppcAxis = ppcAxis or 2
        self.bound = 3
        self.materials: list[dict] = []
        self._material_index: dict[str, int] = {}
        self.colliders: list[dict] = []
        self.destroyed = False
        self.numNodes = self.dims[0] * self.dims[1] * self.dims[2]
        cfg = _Config((ctypes.c_int32 * 3)(*self.dims), self.maxParticles, self.ppcAxis, float(self.bound),
                      self._sortEvery, 0)
        h = ctypes.c_void_p()
        _check(load_library().mpm_sim_create(ctx._h, ctypes.byref(cfg), ctypes.byref(h)), "MPM")
        self._h = h
        self.P = _ParticleView(self)   # GArray-like views (offsets and counts in vec4 units)
        self.grid = _GridView(self)

    # ---- fields ----
    @property
    def sortEvery(self) -> int:
        """Sort every n substeps (mpm.js:101, 591); setting it applies from the next substep on."""
        return self._sortEvery

    @sortEvery.setter
    def sortEvery(self, n: int) -> None:
        self._alive("MPM.sortEvery")
        _check(load_library().mpm_set_sort_every(self._h, int(n)), "MPM.sortEvery")
        self._sortEvery = int(n)

    @property
    def count(self) -> int:
        n = ctypes.c_uint32()
        _check(load_library().mpm_get_count(self._h, ctypes.byref(n)))
        return n.value

    @property
    def stepCount(self) -> int:
        n = ctypes.c_uint64()
        _check(load_library().mpm_get_step_count(self._h, ctypes.byref(n)))
        return n.value

    @stepCount.setter
    def stepCount(self, n: int) -> None:
        _check(load_library().mpm_set_step_count(self._h, int(n)))

    @property
    def substepDt(self) -> float:  # mpm.js:130
        return (1 / 60) * self.timeScale / self.substeps

    @property
    def substeps(self) -> int:
        return self._substeps

    @substeps.setter
    def substeps(self, n) -> None:  # mpm.js:137
        self._retime(lambda: setattr(self, "_substeps", max(1, round(n))))

    @property
    def timeScale(self) -> float:
        return self._timeScale

    @timeScale.setter
    def timeScale(self, x) -> None:
        self._retime(lambda: setattr(self, "_timeScale", max(1e-6, float(x))))

    def _retime(self, change) -> None:  # mpm.js:140-145
        before = self.substepDt
        change()
        k = self.substepDt / before
        if self.count and abs(k - 1) > 1e-12:
            _check(load_library().mpm_rescale(self._h, k), "MPM.retime")

    # ---- materials (mpm.js:148-199) ----
    def _material(self, name_or_def) -> int:
        if isinstance(name_or_def, str):
            if name_or_def not in MATERIALS:
                raise SIError(f"MPM: unknown material {json.dumps(name_or_def)} (use {', '.join(MATERIALS)} or a "
                              "definition with a kind)", "invalid-argument")
            d = {**MATERIALS[name_or_def], "name": name_or_def}
        else:
            d = dict(name_or_def or {})
        if d.get("kind") is None:
            raise SIError(f"MPM: unknown material {json.dumps(name_or_def)}", "invalid-argument")
        if isinstance(d["kind"], str):
            if d["kind"] not in MPM_KINDS:
                raise SIError(f'MPM: unknown material kind "{d["kind"]}" (use {", ".join(MPM_KINDS)})', "invalid-argument")
            d["kind"] = MPM_KINDS[d["kind"]]
        if d["kind"] not in (FLUID, JELLY, SNOW, SAND):
            raise SIError(f"MPM: material kind must be 0–3 (got {d['kind']})", "invalid-argument")
        density = d.get("density") if d.get("density") is not None else 1
        if not (density > 0 and density <= MAX_DENSITY):
            raise SIError(f"MPM: density {density} is outside (0, {MAX_DENSITY}] — densities are relative to water (1), "
                          "not kg/m³", "invalid-argument")
        key = json.dumps(d)
        if key in self._material_index:
            return self._material_index[key]
        if len(self.materials) >= 8:
            raise SIError("MPM: at most 8 distinct materials at once (clear() frees them)", "invalid-argument")
        idx = len(self.materials)
        self.materials.append(d)
        self._material_index[key] = idx
        self._upload_materials()
        return idx

    def material_table(self) -> np.ndarray:
        tbl = np.zeros((8, 12), np.float32)
        for k, m in enumerate(self.materials):
            tbl[k] = material_row(m, self.ppcAxis)
        return tbl

    def _upload_materials(self) -> None:
        tbl = self.material_table()
        _check(load_library().mpm_set_materials(self._h, _ptr(tbl), 8), "MPM materials")

    # ---- emission (mpm.js:207-232) ----
    def add(self, shape: dict, material=None, velocity=None, jitter=None, seed=None, noise: float = 0.0,
            noiseSeed=None) -> int:
        """Fill a shape with ppcAxis^3 particles per cell (mpm.js:207-232). The options may also come as
        the reference's dict, add(shape, {"material": ..., "velocity": ..., "jitter": ..., "seed": ...}).
        noise: std-dev of Gaussian velocity noise (world units/s) — an extension the reference does
        not have."""
        self._alive("MPM.add")
        if isinstance(material, dict) and "kind" not in material:  # the reference's opts object
            o = material
            material, velocity = o.get("material"), o.get("velocity", velocity)
            jitter, seed = o.get("jitter", jitter), o.get("seed", seed)
        mat = self._material(material or shape.get("material") or "water")
        vel = velocity or shape.get("velocity") or [0, 0, 0]
        count = self.count
        if "points" in shape:
            pts = np.asarray(shape["points"], np.float32).reshape(-1, 3)
        else:
            pts = sample_shape(shape, self.dx / self.ppcAxis, 0.6 if jitter is None else jitter,
                               count + 1 if seed is None else seed)
        n = min(len(pts), self.maxParticles - count)
        if n <= 0:
            return 0
        pts = np.ascontiguousarray(pts[:n])
        rec = np.zeros((n, 32), np.float32)
        s = self.substepDt / self.dx
        load_library().mpm_pack_particles(
            _ptr(pts), n, _ptr(np.asarray(self.origin, np.float64)), self.dx, _ptr(np.asarray(self.dims, np.int32)),
            float(self.bound), mat, _ptr(np.asarray(vel, np.float64)), s, float(noise),
            (noiseSeed if noiseSeed is not None else (count + 1) * 2654435761) & 0xFFFFFFFF, _ptr(rec))
        self.write_particles(rec, count)
        _check(load_library().mpm_set_count(self._h, count + n))
        return n

    def clear(self) -> None:  # mpm.js:238-242
        _check(load_library().mpm_set_count(self._h, 0))
        self.materials = []
        self._material_index.clear()

    def setColliders(self, lst: list[dict]) -> None:  # mpm.js:274
        self.colliders = list(lst)

    def collider_table(self) -> np.ndarray:  # mpm.js:276-293
        data = np.zeros(16 * 12, np.float32)
        u = data.view(np.uint32)
        s = self.substepDt / self.dx
        for k, c in enumerate(self.colliders[:16]):
            o = k * 12
            typ = c.get("type") or "sphere"
            u[o] = {"sphere": 0, "box": 1, "plane": 2}[typ]
            ctr = c.get("center") or [0, 0, 0]
            for a in range(3):
                data[o + 1 + a] = (ctr[a] - self.origin[a]) / self.dx
            if typ == "plane":
                nrm = c.get("normal") or [0, 1, 0]
                data[o + 4:o + 7] = nrm
            elif typ == "box":
                h = c.get("half") or [0.1, 0.1, 0.1]
                data[o + 4:o + 7] = [h[0] / self.dx, h[1] / self.dx, h[2] / self.dx]
            else:
                data[o + 4] = (c.get("radius") if c.get("radius") is not None else 0.1) / self.dx
            data[o + 7] = c.get("friction") if c.get("friction") is not None else 0.3
            v = c.get("velocity") or [0, 0, 0]
            data[o + 8:o + 11] = [v[0] * s, v[1] * s, v[2] * s]
        return data

    def lattice_gravity(self) -> np.ndarray:  # mpm.js:588, f64 -> f32
        dt = self.substepDt
        return np.array([(x * dt * dt) / self.dx for x in self.gravity], np.float32)

    # ---- stepping (mpm.js:582-621) ----
    def step(self) -> None:
        if self.destroyed:
            raise SIError("MPM.step: it was destroyed", "destroyed")
        if self.count == 0:
            return
        L = load_library()
        tbl = self.collider_table()
        _check(L.mpm_set_colliders(self._h, _ptr(tbl), min(16, len(self.colliders))), "MPM colliders")
        g = self.lattice_gravity()
        _check(L.mpm_step(self._h, self.substeps, _ptr(g), 0.4, 0.9), "MPM.step")

    def p2g(self, P=None, forceGlobal: bool = False, grid=None, force_global: bool | None = None) -> None:
        """MPM.p2g(P, forceGlobal, grid) (mpm.js:607-612): clear the grid and scatter the particles; the
        grid is left in i32 fixed point. P must be the simulation's own particle buffer; grid is this
        simulation's (default) or a gpu.array('vec4i', numNodes), which receives the result."""
        self._alive("MPM.p2g")
        if P is not None and P is not self.P:
            raise SIError("MPM.p2g: only the simulation's own particle buffer is supported", "unsupported")
        fg = 1 if (forceGlobal if force_global is None else force_global) else 0
        if grid is None or grid is self.grid:
            _check(load_library().mpm_p2g(self._h, fg), "MPM.p2g")
            return
        if not isinstance(grid, GArray) or grid.kind != "vec4i" or grid.length != self.numNodes:
            raise SIError(f"MPM.p2g: grid must be a vec4i array of {self.numNodes} nodes", "invalid-argument")
        _check(load_library().mpm_p2g_to(self._h, fg, _ptr(grid._data)), "MPM.p2g")

    def _sort(self) -> None:
        self._alive("MPM._sort")
        _check(load_library().mpm_sort(self._h), "MPM._sort")

    # ---- data ----
    def write_particles(self, rec: np.ndarray, offset: int = 0) -> None:
        """P.write (records of 32 float32, reference layout)."""
        self._alive("MPM.P.write")
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 32)
        _check(load_library().mpm_write_particles(self._h, _ptr(rec), offset, rec.shape[0]), "MPM.P.write")

    def read_particles(self, n: int | None = None, offset: int = 0) -> np.ndarray:
        """P.read (blocking): (n, 32) float32 records in the reference layout."""
        self._alive("MPM.P.read")
        n = self.count if n is None else n
        out = np.empty((n, 32), np.float32)
        _check(load_library().mpm_read_particles(self._h, _ptr(out), offset, n), "MPM.P.read")
        return out

    def read_particles_async(self, n: int | None = None, offset: int = 0) -> "PendingRead":
        """P.read that resolves later: the snapshot is taken in stream order now, result() waits for
        it (the caller's thread is free meanwhile)."""
        self._alive("MPM.P.read")
        n = self.count if n is None else n
        h = ctypes.c_void_p()
        _check(load_library().mpm_read_particles_begin(self._h, offset, n, ctypes.byref(h)), "MPM.P.read")
        return PendingRead(h, np.empty((n, 32), np.float32), "MPM.P.read")

    def read_grid_async(self) -> "PendingRead":
        """grid.read that resolves later (see read_particles_async)."""
        self._alive("MPM.grid.read")
        h = ctypes.c_void_p()
        _check(load_library().mpm_read_grid_begin(self._h, ctypes.byref(h)), "MPM.grid.read")
        return PendingRead(h, np.empty((self.numNodes, 4), np.int32), "MPM.grid.read")

    def read_grid(self) -> np.ndarray:
        """grid.read (blocking): (numNodes, 4) int32 words in the reference layout."""
        self._alive("MPM.grid.read")
        out = np.empty((self.numNodes, 4), np.int32)
        _check(load_library().mpm_read_grid(self._h, _ptr(out)), "MPM.grid.read")
        return out

    def substep_begin(self) -> None:
        """First part of one substep (sort when due, P2G, grid update): the grid then holds f32
        (v, m) for read_grid / write_grid, as at the reference's coupling call (mpm.js:596)."""
        self._alive("MPM.substep_begin")
        _check(load_library().mpm_substep_begin(self._h, _ptr(self.lattice_gravity()), 0.4), "MPM.substep_begin")

    def substep_end(self) -> None:
        """G2P: finishes the substep begun with substep_begin."""
        self._alive("MPM.substep_end")
        _check(load_library().mpm_substep_end(self._h, 0.9), "MPM.substep_end")

    def write_grid(self, words: np.ndarray) -> None:
        """Writes every node's f32 (v.xyz, mass) bits, reference layout (numNodes, 4) int32."""
        self._alive("MPM.grid.write")
        w = np.ascontiguousarray(words, np.int32).reshape(self.numNodes, 4)
        _check(load_library().mpm_write_grid(self._h, _ptr(w)), "MPM.grid.write")

    def set_grid_hook(self, fn) -> None:
        """Native after-grid-update hook: fn(user, GridView*) is a GRID_HOOK (ctypes) or None."""
        self._hook = fn  # keep the ctypes thunk alive
        _check(load_library().mpm_set_grid_hook(self._h, fn if fn is not None else GRID_HOOK(), None), "MPM.set_grid_hook")

    def read_points(self, velocity: bool = False):
        """World positions + material ids (count, 4) float32 computed on the device (readPositions,
        mpm.js:624-630), and optionally velocities in world units/s with J."""
        self._alive("MPM.read_points")
        n = self.count
        pts = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32) if velocity else None
        _check(load_library().mpm_read_points(self._h, _ptr(pts), _ptr(vel) if velocity else None, 0, n,
                                              _ptr(np.asarray(self.origin, np.float64)), self.dx,
                                              self.dx / self.substepDt), "MPM.read_points")
        return (pts, vel) if velocity else pts

    def particle_streams(self):
        """Device pointers of the 8 particle streams and their capacity (valid until the next step).

        Float f of a reference record is component ``REC_SLOT[f] % 4`` of stream ``REC_SLOT[f] // 4``."""
        ptrs = (ctypes.c_void_p * 8)()
        cap = ctypes.c_uint32()
        _check(load_library().mpm_particle_streams(self._h, ptrs, ctypes.byref(cap)), "MPM.particle_streams")
        return [p or 0 for p in ptrs], cap.value

    def readPositions(self) -> np.ndarray:  # mpm.js:624-630
        raw = self.read_particles()
        return (raw[:, :3].astype(np.float64) * self.dx + np.asarray(self.origin)).astype(np.float32)

    def sync(self) -> None:
        _check(load_library().mpm_sync(self._h))

    def set_grid_law(self, on: bool) -> None:
        """Starts (from zero) or stops the grid update's momentum bookkeeping (mpm_set_grid_law)."""
        _check(load_library().mpm_set_grid_law(self._h, 1 if on else 0), "MPM.set_grid_law")

    def read_grid_law(self, reset: bool = False) -> dict:
        """The bookkeeping sums since they were started or last reset (see grid_law)."""
        self._alive("MPM.read_grid_law")
        return grid_law(self._h, reset=reset)

    TUNE = {"g2p_waves": 2, "p2g_waves": 4, "g2p_items": 5, "fuse": 6, "fuse_shape": 7, "drift": 8, "carry": 9}

    def set_tuning(self, **knobs) -> None:
        """Kernel variant knobs (result-neutral): g2p_waves, p2g_waves, g2p_items, fuse, fuse_shape, drift, carry."""
        for k, v in knobs.items():
            _check(load_library().mpm_set_tuning(self._h, self.TUNE[k], int(v)), "MPM.set_tuning")

    def set_profiling(self, on, every: int = 1) -> None:
        """True / 1: events between the phases of every substep; 2: G2P launches only, timed by
        their own dispatch (no extra packets on the stream), launches 0, every, 2 every, ... (the
        others run without events); False / 0: off."""
        _check(load_library().mpm_set_profiling(self._h, int(on)))
        _check(load_library().mpm_set_profiling_every(self._h, int(every)))

    def kernel_times(self) -> dict:
        """Profiling mode 2: {"g2p": (ms, launches), "g2p_p2g": (ms, launches)} (G2P alone, fused)."""
        ms = np.zeros(2, np.float32)
        n = np.zeros(2, np.uint32)
        _check(load_library().mpm_kernel_times(self._h, _ptr(ms), _ptr(n)), "MPM.kernel_times")
        return {"g2p": (float(ms[0]), int(n[0])), "g2p_p2g": (float(ms[1]), int(n[1]))}

    def phase_times(self) -> dict:
        ms = np.zeros(4, np.float32)
        _check(load_library().mpm_phase_times(self._h, _ptr(ms)))
        return dict(zip(["mpm sort", "mpm p2g", "mpm grid", "mpm g2p"], ms.tolist()))

    def palette(self) -> list:
        return [m.get("color") or [0.7, 0.7, 0.7] for m in self.materials]

    def fluidMask(self) -> int:
        return sum(1 << k for k, d in enumerate(self.materials) if d["kind"] == FLUID)

    def _alive(self, what: str) -> None:
        if self.destroyed:
            raise SIError(f"{what}: it was destroyed", "destroyed")

    def destroy(self) -> None:  # mpm.js:262-267 (idempotent)
        if self.destroyed:
            return
        self.destroyed = True
        load_library().mpm_sim_destroy(self._h)
        self._h = None


class _ParticleView:
    """MPM.P: read(n4) returns n4 vec4 (n4 / 8 records) of the particle buffer as a flat float32 array;
    write(data, offset4) writes records at vec4 offset offset4 (mpm.js:229, 625)."""

    def __init__(self, sim: "MPM"):
        self._sim = sim

    def read(self, n4: int | None = None) -> np.ndarray:
        sim = self._sim
        n = sim.count if n4 is None else int(n4) // 8
        return sim.read_particles(n).reshape(-1)

    def write(self, data, offset4: int = 0) -> None:
        self._sim.write_particles(np.asarray(data, np.float32).reshape(-1, 32), int(offset4) // 8)


class _GridView:
    """MPM.grid: read() returns the grid's int32 words (4 per node, reference layout)."""

    def __init__(self, sim: "MPM"):
        self._sim = sim

    def read(self) -> np.ndarray:
        return self._sim.read_grid().reshape(-1)


def node_blocks(dims) -> int:
    """4x4x4-node blocks a scene of these dims occupies in a pool."""
    return ((dims[0] + 3) // 4) * ((dims[1] + 3) // 4) * ((dims[2] + 3) // 4)


class ScenePool:
    """Many independent MLS-MPM scenes advanced by the same launches (include/mpm.h, mpm_pool_*).

    Each scene keeps its own grid, gravity, materials and particles; its results are those of a
    separate MPM simulation of that scene (src/physics/mpm.js:582-604 per scene)."""

    def __init__(self, ctx: Context, max_particles: int, max_scenes: int, max_node_blocks: int, max_dim: int,
                 sort_every: int = 4):
        self.ctx = ctx
        self.cfg = _PoolConfig(int(max_particles), int(max_scenes), int(max_node_blocks), int(max_dim),
                               int(sort_every), 0)
        h = ctypes.c_void_p()
        _check(load_library().mpm_pool_create(ctx._h, ctypes.byref(self.cfg), ctypes.byref(h)), "ScenePool")
        self._h = h

    @staticmethod
    def bytes_for(max_particles: int, max_scenes: int, max_node_blocks: int, max_dim: int) -> int:
        cfg = _PoolConfig(int(max_particles), int(max_scenes), int(max_node_blocks), int(max_dim), 4, 0)
        return int(load_library().mpm_pool_bytes(ctypes.byref(cfg)))

    def _alive(self, what: str) -> None:
        if not self._h:
            raise SIError(f"{what}: pool destroyed", "destroyed")

    def add_scene(self, dims, materials: np.ndarray, records: np.ndarray, gravity, ppc_axis: int = 2,
                  bound: float = 3.0) -> int:
        """Adds a scene (records in the reference layout with scene-local material ids) and returns its id."""
        self._alive("ScenePool.add_scene")
        mats = np.ascontiguousarray(materials, np.float32).reshape(-1, 12)
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 32)
        d = _SceneDesc((ctypes.c_int32 * 3)(*[int(x) for x in dims]), int(ppc_axis), float(bound),
                       (ctypes.c_float * 3)(*[float(x) for x in gravity]), len(mats),
                       mats.ctypes.data if len(mats) else None, rec.ctypes.data if len(rec) else None, len(rec))
        sid = ctypes.c_uint32()
        _check(load_library().mpm_pool_add_scene(self._h, ctypes.byref(d), ctypes.byref(sid)), "ScenePool.add_scene")
        return sid.value

    def _desc(self, dims, materials, records, gravity, ppc_axis, bound):
        mats = np.ascontiguousarray(materials, np.float32).reshape(-1, 12)
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 32)
        d = _SceneDesc((ctypes.c_int32 * 3)(*[int(x) for x in dims]), int(ppc_axis), float(bound),
                       (ctypes.c_float * 3)(*[float(x) for x in gravity]), len(mats),
                       mats.ctypes.data if len(mats) else None, rec.ctypes.data if len(rec) else None, len(rec))
        return d, (mats, rec)

    def add_template(self, dims, materials: np.ndarray, records: np.ndarray, gravity, ppc_axis: int = 2,
                     bound: float = 3.0) -> int:
        """Keeps a scene on the device for repeated admission; returns the template id."""
        self._alive("ScenePool.add_template")
        d, keep = self._desc(dims, materials, records, gravity, ppc_axis, bound)
        tid = ctypes.c_uint32()
        _check(load_library().mpm_pool_add_template(self._h, ctypes.byref(d), ctypes.byref(tid)),
               "ScenePool.add_template")
        return tid.value

    def add_from_template(self, tmpl: int) -> int:
        self._alive("ScenePool.add_from_template")
        sid = ctypes.c_uint32()
        _check(load_library().mpm_pool_add_scene_from_template(self._h, int(tmpl), ctypes.byref(sid)),
               "ScenePool.add_from_template")
        return sid.value

    def autotune(self, trial_substeps: int = 8) -> dict:
        """Times the kernel shapes on the live pool (it advances by the trials) and keeps the best."""
        self._alive("ScenePool.autotune")
        g, p = ctypes.c_int32(), ctypes.c_int32()
        _check(load_library().mpm_pool_autotune(self._h, int(trial_substeps), ctypes.byref(g), ctypes.byref(p)),
               "ScenePool.autotune")
        return self.tuning()

    def tuning(self) -> dict:
        """The current value of every tuning knob."""
        out = {}
        for k, code in MPM.TUNE.items():
            v = ctypes.c_int32()
            _check(load_library().mpm_pool_tuning(self._h, code, ctypes.byref(v)), "ScenePool.tuning")
            out[k] = v.value
        return out

    def remove_scene(self, scene: int) -> None:
        self._alive("ScenePool.remove_scene")
        _check(load_library().mpm_pool_remove_scene(self._h, int(scene)), "ScenePool.remove_scene")

    def step(self, substeps: int, floor_friction: float = 0.4, vmax: float = 0.9) -> None:
        self._alive("ScenePool.step")
        _check(load_library().mpm_pool_step(self._h, int(substeps), float(floor_friction), float(vmax)),
               "ScenePool.step")

    def scene_count(self, scene: int) -> int:
        n = ctypes.c_uint32()
        _check(load_library().mpm_pool_scene_count(self._h, int(scene), ctypes.byref(n)), "ScenePool.scene_count")
        return n.value

    def read_scene(self, scene: int) -> np.ndarray:
        self._alive("ScenePool.read_scene")
        n = self.scene_count(scene)
        out = np.zeros((n, 32), np.float32)
        got = ctypes.c_uint32()
        _check(load_library().mpm_pool_read_scene(self._h, int(scene), _ptr(out), n, ctypes.byref(got)),
               "ScenePool.read_scene")
        return out

    def info(self) -> dict:
        a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(load_library().mpm_pool_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"particles": a.value, "scenes": b.value, "free_node_blocks": c.value}

    def sync(self) -> None:
        _check(load_library().mpm_pool_sync(self._h))

    def use_graphs(self, on: bool) -> None:
        _check(load_library().mpm_pool_use_graphs(self._h, 1 if on else 0))

    def set_tuning(self, **knobs) -> None:
        for k, v in knobs.items():
            _check(load_library().mpm_pool_set_tuning(self._h, MPM.TUNE[k], int(v)), "ScenePool.set_tuning")

    def destroy(self) -> None:
        if self._h:
            load_library().mpm_pool_destroy(self._h)
            self._h = None


class Server:
    """The native serving loop over a ScenePool (include/mpm.h, mpm_server_*). While it runs, its
    thread owns the pool."""

    def __init__(self, pool: ScenePool, substeps: int = 4, floor_friction: float = 0.4, vmax: float = 0.9,
                 graph_after: int = 0):
        self.pool = pool
        cfg = _ServerConfig(int(substeps), float(floor_friction), float(vmax), int(graph_after))
        self._results: dict = {}  # result arrays the native loop writes into, kept alive until settled
        h = ctypes.c_void_p()
        _check(load_library().mpm_server_create(pool._h, ctypes.byref(cfg), ctypes.byref(h)), "Server")
        self._h = h

    def submit(self, tmpl: int, lifetime_steps: int, result: np.ndarray | None = None) -> int:
        """Queues a request; with `result` ((n, 32) float32, kept alive here until wait() reports the
        ticket done) the scene's final records are copied into it before the scene leaves."""
        t = ctypes.c_uint64()
        L = load_library()
        if result is None:
            _check(L.mpm_server_submit(self._h, int(tmpl), int(lifetime_steps), ctypes.byref(t)), "Server.submit")
        else:
            if result.dtype != np.float32 or result.ndim != 2 or result.shape[1] != 32 or not result.flags.c_contiguous:
                raise SIError("Server.submit: result must be a C-contiguous (n, 32) float32 array", "invalid-argument")
            _check(L.mpm_server_submit_result(self._h, int(tmpl), int(lifetime_steps), _ptr(result), result.shape[0],
                                              ctypes.byref(t)), "Server.submit")
            self._results[t.value] = result
        return t.value

    STATES = ("waiting", "running", "done", "rejected")

    def wait(self, ticket: int, timeout_ms: int = -1) -> tuple[str, int]:
        """(state, records copied) of a ticket, after waiting up to timeout_ms (-1: until it settles)."""
        st, n = ctypes.c_int32(), ctypes.c_uint32()
        _check(load_library().mpm_server_wait(self._h, int(ticket), int(timeout_ms), ctypes.byref(st), ctypes.byref(n)),
               "Server.wait")
        state = self.STATES[st.value]
        if state in ("done", "rejected"):
            self._results.pop(int(ticket), None)
        return state, n.value

    def stats(self) -> dict:
        c = _ServerCounters()
        _check(load_library().mpm_server_stats(self._h, ctypes.byref(c)), "Server.stats")
        return {k: getattr(c, k) for k, _ in _ServerCounters._fields_}

    def latencies(self, kind: str) -> np.ndarray:
        code = {"step_ms": 0, "request_s": 1, "wait_s": 2, "age_s": 3, "scene_step_s": 4, "scene_step_age_s": 5}[kind]
        n = ctypes.c_uint64()
        L = load_library()
        _check(L.mpm_server_latencies(self._h, code, None, 0, ctypes.byref(n)), "Server.latencies")
        out = np.zeros(n.value, np.float64)
        if n.value:
            _check(L.mpm_server_latencies(self._h, code, _ptr(out), n.value, ctypes.byref(n)), "Server.latencies")
        return out[:n.value]

    def stop(self, drain: bool = False) -> None:
        if self._h:
            _check(load_library().mpm_server_stop(self._h, 1 if drain else 0), "Server.stop")

    def destroy(self) -> None:
        if self._h:
            load_library().mpm_server_destroy(self._h)
            self._h = None
