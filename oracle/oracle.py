"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference step.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to compare the engine against.  The engine
never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class Config(ctypes.Structure):
    """SimConfig, common.cuh:65-79."""

    _fields_ = [
        ("SIZE_X", ctypes.c_uint32), ("SIZE_Y", ctypes.c_uint32), ("SIZE_Z", ctypes.c_uint32),
        ("PADDING", ctypes.c_uint32), ("GRAVITY", ctypes.c_float), ("DT", ctypes.c_float),
        ("GAMMA", ctypes.c_float), ("COMPRESSION", ctypes.c_float), ("RESOLUTION", ctypes.c_float),
        ("SUBSTEPS", ctypes.c_uint32), ("SDF_RESOLUTION", ctypes.c_uint32), ("GRID_SIZE", ctypes.c_uint32),
        ("GRID_NUMBER", ctypes.c_uint32), ("GRID_BLOCKS", ctypes.c_uint32),
    ]


def build() -> Path:
    subprocess.run(["bash", str(HERE / "build.sh")], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        f, d = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
        i32 = ctypes.POINTER(ctypes.c_int32)
        cp = ctypes.POINTER(Config)
        sigs = {
            "oracle_max_threads": (ctypes.c_int, []),
            "oracle_spline_f32": (ctypes.c_float, [ctypes.c_float]),
            "oracle_spline_f64": (ctypes.c_double, [ctypes.c_double]),
            "oracle_expand_bits": (ctypes.c_uint32, [ctypes.c_uint32]),
            "oracle_morton": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
            "oracle_morton_keys": (None, [ctypes.c_int, f, f, f, i32]),
            "oracle_sort_perm": (None, [ctypes.c_int, i32, i32]),
            "oracle_init_state_f32": (None, [cp, ctypes.c_int, f]),
            "oracle_init_state_f64": (None, [cp, ctypes.c_int, d]),
            "oracle_steps_f32": (None, [cp, ctypes.c_int, f, f, ctypes.c_int, ctypes.c_int]),
            "oracle_steps_f64": (None, [cp, ctypes.c_int, d, d, ctypes.c_int, ctypes.c_int]),
            "oracle_p2g_f32": (None, [cp, ctypes.c_int, f, i32, f]),
            "oracle_p2g_f64": (None, [cp, ctypes.c_int, d, i32, d]),
            "oracle_grid_update_f32": (None, [cp, f, f]),
            "oracle_grid_update_f64": (None, [cp, d, d]),
            "oracle_g2p_f32": (None, [cp, ctypes.c_int, f, f, f]),
            "oracle_g2p_f64": (None, [cp, ctypes.c_int, d, d, d]),
            "oracle_get_sdf_f32": (ctypes.c_float, [cp, ctypes.c_float, ctypes.c_float, ctypes.c_float, f]),
            "oracle_solid_angle_f32": (ctypes.c_float, [f, f]),
            "oracle_point_tri_dist2_f32": (ctypes.c_float, [f, f]),
            "oracle_point_tri_dist2_f64": (ctypes.c_double, [d, d]),
            "oracle_make_sdf_f32": (None, [cp, f, ctypes.c_int, f, ctypes.c_int]),
            "oracle_time_steps_f32": (ctypes.c_double, [cp, ctypes.c_int, f, f, ctypes.c_int, ctypes.c_int]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def to_config(cfg) -> Config:
    """Copies any object with the SimConfig field names into an oracle Config."""
    c = Config()
    for name, _ in Config._fields_:
        setattr(c, name, getattr(cfg, name))
    return c


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def nodes(cfg) -> int:
    return int(cfg.SIZE_X) * int(cfg.SIZE_Y) * int(cfg.SIZE_Z)


def init_state(cfg, parts: np.ndarray, dtype=np.float32) -> np.ndarray:
    """initFMatrices semantics on a copy of a (26, n) buffer."""
    p = np.ascontiguousarray(parts, dtype=dtype).copy()
    n = p.shape[1]
    c = to_config(cfg)
    if dtype == np.float32:
        lib().oracle_init_state_f32(ctypes.byref(c), n, _f(p))
    else:
        lib().oracle_init_state_f64(ctypes.byref(c), n, _d(p))
    return p


def steps(cfg, state: np.ndarray, sdf: np.ndarray, k: int, nthreads: int = 1) -> np.ndarray:
    """k reference substeps on a copy of an initialised (26, n) state."""
    c = to_config(cfg)
    n = state.shape[1]
    if state.dtype == np.float64:
        p = np.ascontiguousarray(state, dtype=np.float64).copy()
        s = np.ascontiguousarray(sdf, dtype=np.float64)
        lib().oracle_steps_f64(ctypes.byref(c), n, _d(p), _d(s), k, nthreads)
    else:
        p = np.ascontiguousarray(state, dtype=np.float32).copy()
        s = np.ascontiguousarray(sdf, dtype=np.float32)
        lib().oracle_steps_f32(ctypes.byref(c), n, _f(p), _f(s), k, nthreads)
    return p


def time_steps(cfg, state: np.ndarray, sdf: np.ndarray, k: int, nthreads: int) -> tuple[float, np.ndarray]:
    c = to_config(cfg)
    p = np.ascontiguousarray(state, dtype=np.float32).copy()
    s = np.ascontiguousarray(sdf, dtype=np.float32)
    t = lib().oracle_time_steps_f32(ctypes.byref(c), p.shape[1], _f(p), _f(s), k, nthreads)
    return t, p


def morton_keys(pos: np.ndarray) -> np.ndarray:
    """setKeys over (3, n) float32 positions."""
    p = np.ascontiguousarray(pos, dtype=np.float32)
    n = p.shape[1]
    keys = np.empty(n, dtype=np.int32)
    x, y, z = (np.ascontiguousarray(p[i]) for i in range(3))
    lib().oracle_morton_keys(n, _f(x), _f(y), _f(z), _i(keys))
    return keys


def sort_perm(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.int32)
    perm = np.empty(k.shape[0], dtype=np.int32)
    lib().oracle_sort_perm(k.shape[0], _i(k), _i(perm))
    return perm


def p2g(cfg, state: np.ndarray, perm=None) -> np.ndarray:
    c = to_config(cfg)
    n = state.shape[1]
    if state.dtype == np.float64:
        p = np.ascontiguousarray(state)
        g = np.zeros((4, nodes(cfg)), dtype=np.float64)
        lib().oracle_p2g_f64(ctypes.byref(c), n, _d(p), None if perm is None else _i(np.ascontiguousarray(perm, np.int32)), _d(g))
    else:
        p = np.ascontiguousarray(state, dtype=np.float32)
        g = np.zeros((4, nodes(cfg)), dtype=np.float32)
        lib().oracle_p2g_f32(ctypes.byref(c), n, _f(p), None if perm is None else _i(np.ascontiguousarray(perm, np.int32)), _f(g))
    return g


def grid_update(cfg, grid4: np.ndarray, sdf: np.ndarray) -> np.ndarray:
    c = to_config(cfg)
    if grid4.dtype == np.float64:
        g = np.ascontiguousarray(grid4, dtype=np.float64).copy()
        lib().oracle_grid_update_f64(ctypes.byref(c), _d(g), _d(np.ascontiguousarray(sdf, dtype=np.float64)))
    else:
        g = np.ascontiguousarray(grid4, dtype=np.float32).copy()
        lib().oracle_grid_update_f32(ctypes.byref(c), _f(g), _f(np.ascontiguousarray(sdf, dtype=np.float32)))
    return g


def g2p(cfg, state: np.ndarray, grid4: np.ndarray, sdf: np.ndarray) -> np.ndarray:
    c = to_config(cfg)
    n = state.shape[1]
    if state.dtype == np.float64:
        p = np.ascontiguousarray(state, dtype=np.float64).copy()
        lib().oracle_g2p_f64(ctypes.byref(c), n, _d(p), _d(np.ascontiguousarray(grid4, np.float64)),
                             _d(np.ascontiguousarray(sdf, np.float64)))
    else:
        p = np.ascontiguousarray(state, dtype=np.float32).copy()
        lib().oracle_g2p_f32(ctypes.byref(c), n, _f(p), _f(np.ascontiguousarray(grid4, np.float32)),
                             _f(np.ascontiguousarray(sdf, np.float32)))
    return p


def make_sdf(cfg, tris: np.ndarray, nthreads: int = 1) -> np.ndarray:
    c = to_config(cfg)
    t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    out = np.empty(nodes(cfg), dtype=np.float32)
    lib().oracle_make_sdf_f32(ctypes.byref(c), _f(t) if t.size else None, t.shape[0], _f(out), nthreads)
    return out


def spline(x: float, dtype=np.float32) -> float:
    if dtype == np.float64:
        return lib().oracle_spline_f64(x)
    return lib().oracle_spline_f32(x)


def solid_angle(P, tri) -> float:
    return lib().oracle_solid_angle_f32(_f(np.ascontiguousarray(P, np.float32)), _f(np.ascontiguousarray(tri, np.float32)))


def point_tri_dist2(P, tri, dtype=np.float32) -> float:
    if dtype == np.float64:
        return lib().oracle_point_tri_dist2_f64(_d(np.ascontiguousarray(P, np.float64)),
                                                _d(np.ascontiguousarray(tri, np.float64)))
    return lib().oracle_point_tri_dist2_f32(_f(np.ascontiguousarray(P, np.float32)),
                                            _f(np.ascontiguousarray(tri, np.float32)))


def get_sdf(cfg, p, sdf) -> float:
    c = to_config(cfg)
    return lib().oracle_get_sdf_f32(ctypes.byref(c), float(p[0]), float(p[1]), float(p[2]),
                                    _f(np.ascontiguousarray(sdf, np.float32)))
