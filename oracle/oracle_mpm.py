"""ctypes wrapper of the CPU oracle (liboracle_mpm.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this module,
and only as the checker / CPU reference timing. The product library never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle_mpm.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(LIB)
    vp, u32, i32, u64, f32, f64, i64 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64,
                                        ctypes.c_float, ctypes.c_double, ctypes.c_int64)
    sig = {
        "or_expf": (f32, [f32]), "or_logf": (f32, [f32]), "or_morton3": (u32, [u32, u32, u32]),
        "or_to_fixed": (i32, [f32]), "or_rsqrt_1to2": (None, [vp, vp, u64]), "or_bspline": (None, [vp, vp]), "or_svd3": (None, [vp, vp, vp, vp]),
        "or_scan_u32": (None, [vp, vp, u64]), "or_counting_sort": (ctypes.c_int, [vp, u64, u32, vp, vp]),
        "or_material_row": (None, [ctypes.c_int, f64, f64, f64, f64, f64, f64, f64, f64, ctypes.c_int, vp]),
        "or_sample_shape": (i64, [ctypes.c_int, vp, f64, f64, u32, vp, i64]),
        "or_pack_particles": (None, [vp, i64, vp, f64, vp, f64, u32, vp, f64, f64, u32, vp]),
        "or_create": (vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, u32]),
        "or_destroy": (None, [vp]), "or_set_materials": (None, [vp, vp, u32]),
        "or_set_colliders": (None, [vp, vp, u32]), "or_write_particles": (None, [vp, vp, u32, u32]),
        "or_read_particles": (None, [vp, vp, u32, u32]), "or_set_count": (None, [vp, u32]),
        "or_set_step_count": (None, [vp, u64]), "or_get_step_count": (u64, [vp]),
        "or_set_sort_every": (None, [vp, ctypes.c_int]),
        "or_read_grid": (None, [vp, vp]), "or_write_grid": (None, [vp, vp]), "or_p2g": (None, [vp]),
        "or_grid_update": (None, [vp, vp, f32]), "or_g2p": (None, [vp, f32]), "or_sort": (None, [vp]),
        "or_rescale": (None, [vp, f32]), "or_keys": (None, [vp, vp]), "or_step": (None, [vp, u32, vp, f32, f32]),
        "or_num_threads": (ctypes.c_int, []), "or_set_num_threads": (None, [ctypes.c_int]),
        "or_set_grid_law": (None, [vp, ctypes.c_int]), "or_read_grid_law": (None, [vp, vp, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray) -> ctypes.c_void_p:
    # data_as keeps a reference to the array, so temporaries outlive the call they are passed to
    return a.ctypes.data_as(ctypes.c_void_p)


def expf(x: float) -> float:
    return lib().or_expf(x)


def logf(x: float) -> float:
    return lib().or_logf(x)


def rsqrt_1to2(x: np.ndarray) -> np.ndarray:
    """The oracle's inverseSqrt on [1, 2], elementwise (float32)."""
    x = np.ascontiguousarray(np.asarray(x, np.float32))
    y = np.empty_like(x)
    lib().or_rsqrt_1to2(_p(x), _p(y), x.size)
    return y


def svd3(F: np.ndarray):
    """F (3x3, float32, row-major numpy) -> U, s, V (numpy row-major)."""
    Fc = np.ascontiguousarray(np.asarray(F, np.float32).T)  # column-major storage
    U, s, V = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
    lib().or_svd3(_p(Fc), _p(U), _p(s), _p(V))
    return U.reshape(3, 3).T.copy(), s, V.reshape(3, 3).T.copy()


def scan(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib().or_scan_u32(_p(a), _p(out), a.size)
    return out


def counting_sort(keys: np.ndarray, nb: int):
    keys = np.ascontiguousarray(keys, np.uint32)
    order = np.empty(keys.size, np.uint32)
    start = np.empty(nb + 1, np.uint32)
    if lib().or_counting_sort(_p(keys), keys.size, nb, _p(order), _p(start)) != 0:
        raise ValueError("key out of range")
    return order, start


def material_row(m: dict, ppc_axis: int = 2) -> np.ndarray:
    g = lambda k, d: float(m[k]) if m.get(k) is not None else d  # noqa: E731
    out = np.zeros(12, np.float32)
    lib().or_material_row(int(m["kind"]), g("density", 1.0), g("stiffness", 1.0), g("poisson", 0.3),
                          g("viscosity", 0.0), g("hardening", 10.0), g("compression", 0.025), g("stretch", 0.0075),
                          g("friction", 33.0), ppc_axis, _p(out))
    return out


def sample_shape(kind: int, prm, spacing: float, jitter: float, seed: int) -> np.ndarray:
    p = np.asarray(list(prm) + [0.0] * (6 - len(prm)), np.float64)
    n = lib().or_sample_shape(kind, _p(p), spacing, jitter, seed & 0xFFFFFFFF, None, 0)
    out = np.empty((n, 3), np.float32)
    lib().or_sample_shape(kind, _p(p), spacing, jitter, seed & 0xFFFFFFFF, _p(out), n)
    return out


def pack_particles(pts, origin, dx, dims, bound, mat, vel, s, noise_sigma=0.0, noise_seed=1) -> np.ndarray:
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    rec = np.zeros((pts.shape[0], 32), np.float32)
    lib().or_pack_particles(_p(pts), pts.shape[0], _p(np.asarray(origin, np.float64)), dx,
                            _p(np.asarray(dims, np.int32)), bound, mat, _p(np.asarray(vel, np.float64)), s,
                            noise_sigma, noise_seed & 0xFFFFFFFF, _p(rec))
    return rec


class Scene:
    """The reference's MPM state on the CPU (reference layouts: 32-float records, x-fastest grid)."""

    def __init__(self, dims, cap: int, ppc_axis: int = 2, sort_every: int = 4):
        self.dims = list(dims)
        self.cap = cap
        self._h = lib().or_create(dims[0], dims[1], dims[2], ppc_axis, sort_every, cap)
        self.nodes = dims[0] * dims[1] * dims[2]

    def set_materials(self, tbl: np.ndarray) -> None:
        tbl = np.ascontiguousarray(tbl, np.float32).reshape(-1)
        lib().or_set_materials(self._h, _p(tbl), tbl.size // 12)

    def set_colliders(self, tbl: np.ndarray, n: int) -> None:
        tbl = np.ascontiguousarray(tbl, np.float32).reshape(-1)
        lib().or_set_colliders(self._h, _p(tbl), n)

    def write(self, rec: np.ndarray, offset: int = 0) -> None:
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 32)
        lib().or_write_particles(self._h, _p(rec), offset, rec.shape[0])

    def read(self, n: int, offset: int = 0) -> np.ndarray:
        out = np.empty((n, 32), np.float32)
        lib().or_read_particles(self._h, _p(out), offset, n)
        return out

    def set_count(self, n: int) -> None:
        lib().or_set_count(self._h, n)

    @property
    def step_count(self) -> int:
        return lib().or_get_step_count(self._h)

    @step_count.setter
    def step_count(self, n: int) -> None:
        lib().or_set_step_count(self._h, n)

    def set_sort_every(self, n: int) -> None:
        lib().or_set_sort_every(self._h, n)

    def grid(self) -> np.ndarray:
        out = np.empty((self.nodes, 4), np.int32)
        lib().or_read_grid(self._h, _p(out))
        return out

    def write_grid(self, words: np.ndarray) -> None:
        words = np.ascontiguousarray(words, np.int32)
        lib().or_write_grid(self._h, _p(words))

    def p2g(self) -> None:
        lib().or_p2g(self._h)

    def grid_update(self, g, floor_friction: float = 0.4) -> None:
        lib().or_grid_update(self._h, _p(np.asarray(g, np.float32)), floor_friction)

    def g2p(self, vmax: float = 0.9) -> None:
        lib().or_g2p(self._h, vmax)

    def sort(self) -> None:
        lib().or_sort(self._h)

    def keys(self, n: int) -> np.ndarray:
        out = np.empty(n, np.uint32)
        lib().or_keys(self._h, _p(out))
        return out

    def rescale(self, k: float) -> None:
        lib().or_rescale(self._h, k)

    def set_grid_law(self, on: bool) -> None:
        """Starts (from zero) or stops the grid update's momentum bookkeeping."""
        lib().or_set_grid_law(self._h, 1 if on else 0)

    def read_grid_law(self, reset: bool = False) -> np.ndarray:
        """The bookkeeping sums since the last reset as 16 uint64 (see law_node in mpm_oracle.cpp)."""
        out = np.zeros(16, np.uint64)
        lib().or_read_grid_law(self._h, _p(out), 1 if reset else 0)
        return out

    def step(self, substeps: int, g, floor_friction: float = 0.4, vmax: float = 0.9) -> None:
        lib().or_step(self._h, substeps, _p(np.asarray(g, np.float32)), floor_friction, vmax)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_destroy(self._h)
            self._h = None
