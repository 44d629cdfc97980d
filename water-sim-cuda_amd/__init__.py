"""Python host mirror of the engine's C ABI (libmpm.so, include/mpm.h).

Mirrors the reference's Engine interface (engine.h:14-43) with the same
method names -- ``init``, ``step``, ``initSDF``, ``getNumber``,
``getParticles``, ``getPositions``, ``getGrid``, ``sortParticles`` -- over
ctypes, with numpy arrays at the boundary.  The compute path is the HIP
library only: importing this package never falls back to a CPU
implementation, and constructing an Engine without a GPU or without the
built library raises.

The directory name contains hyphens; import it with
``importlib.import_module("water-sim-cuda_amd")`` after putting the
repository root on ``sys.path``.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmpm.so"

PARTICLE_ATTRS = 26
ATTR_NAMES = (
    ["pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z"]
    + [f"C{r}{c}" for r in range(3) for c in range(3)]
    + [f"F{r}{c}" for r in range(3) for c in range(3)]
    + ["vol", "mass"]
)

MPM_OK = 0
MPM_ELIMIT = -6
SORT_NOALLOC_MAX = 1 << 21  # mpm_sort_keys_noalloc's fixed capacity (MPM_SORT_NOALLOC_MAX)


class MPMError(RuntimeError):
    pass


class SimConfig(ctypes.Structure):
    """SimConfig, common.cuh:65-79 (same fields and order)."""

    _fields_ = [
        ("SIZE_X", ctypes.c_uint32),
        ("SIZE_Y", ctypes.c_uint32),
        ("SIZE_Z", ctypes.c_uint32),
        ("PADDING", ctypes.c_uint32),
        ("GRAVITY", ctypes.c_float),
        ("DT", ctypes.c_float),
        ("GAMMA", ctypes.c_float),
        ("COMPRESSION", ctypes.c_float),
        ("RESOLUTION", ctypes.c_float),
        ("SUBSTEPS", ctypes.c_uint32),
        ("SDF_RESOLUTION", ctypes.c_uint32),
        ("GRID_SIZE", ctypes.c_uint32),
        ("GRID_NUMBER", ctypes.c_uint32),
        ("GRID_BLOCKS", ctypes.c_uint32),
    ]

    def grid_shape(self) -> tuple[int, int, int]:
        """(SIZE_Z, SIZE_Y, SIZE_X): node arrays reshape to this (x fastest)."""
        return (int(self.SIZE_Z), int(self.SIZE_Y), int(self.SIZE_X))

    @property
    def nodes(self) -> int:
        return int(self.SIZE_X) * int(self.SIZE_Y) * int(self.SIZE_Z)

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ConfigObject(ctypes.Structure):
    _fields_ = [
        ("model_path", ctypes.c_char * 256),
        ("scale", ctypes.c_float),
        ("x", ctypes.c_float),
        ("y", ctypes.c_float),
        ("z", ctypes.c_float),
    ]


_lib = None


def library_path() -> Path:
    return LIB_PATH


def load_library() -> ctypes.CDLL:
    """Loads libmpm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise MPMError(f"{LIB_PATH} is missing: build it first (python water-sim-cuda_amd/build.py)")
    lib = ctypes.CDLL(str(LIB_PATH))
    c_float_p = ctypes.POINTER(ctypes.c_float)
    c_int_p = ctypes.POINTER(ctypes.c_int)
    c_i32_p = ctypes.POINTER(ctypes.c_int32)
    c_i64_p = ctypes.POINTER(ctypes.c_int64)
    c_dbl_p = ctypes.POINTER(ctypes.c_double)
    eng = ctypes.c_void_p
    sigs = {
        "mpm_last_error": (ctypes.c_char_p, []),
        "mpm_version": (ctypes.c_int, []),
        "mpm_config_derive": (ctypes.c_int, [ctypes.POINTER(SimConfig)]),
        "mpm_config_parse_file": (
            ctypes.c_int,
            [ctypes.c_char_p, ctypes.POINTER(SimConfig), ctypes.POINTER(ConfigObject), c_int_p,
             ctypes.POINTER(ConfigObject), c_int_p, ctypes.c_int],
        ),
        "mpm_config_parse_string": (
            ctypes.c_int,
            [ctypes.c_char_p, ctypes.POINTER(SimConfig), ctypes.POINTER(ConfigObject), c_int_p,
             ctypes.POINTER(ConfigObject), c_int_p, ctypes.c_int],
        ),
        "mpm_config_get_int": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, c_i64_p]),
        "mpm_engine_create": (ctypes.c_int, [ctypes.POINTER(SimConfig), ctypes.POINTER(ctypes.c_void_p)]),
        "mpm_engine_destroy": (ctypes.c_int, [eng]),
        "mpm_engine_set_option": (ctypes.c_int, [eng, ctypes.c_char_p, ctypes.c_int64]),
        "mpm_engine_get_option": (ctypes.c_int, [eng, ctypes.c_char_p, c_i64_p]),
        "mpm_engine_init_particles": (ctypes.c_int, [eng, ctypes.c_int, c_float_p]),
        "mpm_engine_load_state": (ctypes.c_int, [eng, ctypes.c_int, c_float_p]),
        "mpm_engine_init_sdf_from_triangles": (ctypes.c_int, [eng, c_float_p, ctypes.c_int]),
        "mpm_engine_set_sdf": (ctypes.c_int, [eng, c_float_p]),
        "mpm_engine_get_sdf": (ctypes.c_int, [eng, c_float_p]),
        "mpm_engine_step": (ctypes.c_int, [eng]),
        "mpm_engine_steps": (ctypes.c_int, [eng, ctypes.c_int]),
        "mpm_engine_synchronize": (ctypes.c_int, [eng]),
        "mpm_engine_get_number": (ctypes.c_int, [eng]),
        "mpm_engine_get_particles": (ctypes.c_int, [eng, c_float_p]),
        "mpm_engine_get_positions": (ctypes.c_int, [eng, c_float_p]),
        "mpm_engine_get_sort": (ctypes.c_int, [eng, c_i32_p, c_i32_p]),
        "mpm_engine_get_grid": (ctypes.c_int, [eng, c_float_p]),
        "mpm_engine_get_stage_times": (ctypes.c_int, [eng, c_dbl_p, ctypes.c_int]),
        "mpm_engine_get_counters": (ctypes.c_int, [eng, c_i64_p, ctypes.c_int]),
        "mpm_engine_debug_p2g": (ctypes.c_int, [eng, c_float_p]),
        "mpm_engine_debug_grid_update": (ctypes.c_int, [eng, c_float_p, c_float_p]),
        "mpm_engine_debug_g2p": (ctypes.c_int, [eng, c_float_p]),
        "mpm_engine_set_dt_scale": (ctypes.c_int, [eng, ctypes.c_float]),
        "mpm_engine_check": (ctypes.c_int, [eng, ctypes.c_float, c_dbl_p]),
        "mpm_engine_set_corrector": (ctypes.c_int, [eng, ctypes.c_void_p]),
        "mpm_engine_collect_pairs": (ctypes.c_int, [eng, ctypes.c_int, c_float_p, c_float_p, ctypes.c_int, c_int_p]),
        "mpm_engine_debug_features": (ctypes.c_int, [eng, c_i32_p, c_float_p, ctypes.c_int, c_int_p]),
        "mpm_corrector_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
        "mpm_corrector_destroy": (ctypes.c_int, [ctypes.c_void_p]),
        "mpm_corrector_set_params": (ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int]),
        "mpm_corrector_get_params": (ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int]),
        "mpm_corrector_set_norm": (ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_float]),
        "mpm_corrector_set_clamp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
        "mpm_corrector_set_deadzone": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
        "mpm_corrector_forward": (ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p]),
        "mpm_corrector_load_dataset": (ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_int]),
        "mpm_corrector_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_float_p, c_float_p]),
        "mpm_corrector_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, c_float_p]),
        "mpm_engine_snapshot": (ctypes.c_int, [eng]),
        "mpm_engine_restore": (ctypes.c_int, [eng]),
        "mpm_batch_create": (ctypes.c_int, [ctypes.POINTER(SimConfig), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
        "mpm_batch_init_particles": (ctypes.c_int, [eng, c_int_p, c_float_p]),
        "mpm_batch_init_sdf_from_triangles": (ctypes.c_int, [eng, ctypes.c_int, c_float_p, ctypes.c_int]),
        "mpm_batch_set_sdf": (ctypes.c_int, [eng, ctypes.c_int, c_float_p]),
        "mpm_batch_get_grid": (ctypes.c_int, [eng, ctypes.c_int, c_float_p]),
        "mpm_batch_scene_info": (ctypes.c_int, [eng, ctypes.c_int, c_i64_p]),
        "mpm_probe_copy": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, c_dbl_p]),
        "mpm_probe_copy_wide": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, c_dbl_p]),
        "mpm_host_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
        "mpm_host_free": (ctypes.c_int, [ctypes.c_void_p]),
        "mpm_device_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
        "mpm_device_free": (ctypes.c_int, [ctypes.c_void_p]),
        "mpm_copy_to_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "mpm_copy_to_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "mpm_sort_keys_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        "mpm_obj_triangles": (ctypes.c_int, [ctypes.c_char_p, c_float_p, ctypes.c_int, c_int_p]),
        "mpm_voxelize_obj": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_float, c_float_p, ctypes.c_int64, c_i64_p]),
        "mpm_obj_water": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_float, ctypes.c_float, c_float_p, ctypes.c_int,
                                         c_float_p, ctypes.c_int64, c_i64_p]),
        "mpm_obj_collider": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_float, ctypes.c_float, c_float_p, c_float_p,
                                            ctypes.c_int, c_int_p]),
        "mpm_scene_from_config": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(SimConfig), c_float_p, ctypes.c_int64,
                                                 c_i64_p, c_float_p, ctypes.c_int, c_int_p]),
        "mpm_scene_fill": (
            ctypes.c_int,
            [ctypes.c_int, c_float_p, c_float_p, ctypes.c_uint64, c_float_p, ctypes.c_int64, c_i64_p],
        ),
        "mpm_mesh_uv_sphere": (ctypes.c_int, [c_float_p, ctypes.c_float, ctypes.c_int, ctypes.c_int, c_float_p, c_int_p]),
        "mpm_mesh_u_channel": (ctypes.c_int, [c_float_p, c_float_p, ctypes.c_float, ctypes.c_float, c_float_p, c_int_p]),
        "mpm_mesh_box": (ctypes.c_int, [c_float_p, c_float_p, c_float_p, c_int_p]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != MPM_OK:
        msg = load_library().mpm_last_error().decode(errors="replace")
        raise MPMError(f"libmpm error {rc}: {msg}")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


# ---- configuration ----

def make_config(**kw) -> SimConfig:
    """SimConfig from keyword arguments (reference key names); GRID_* derived."""
    c = SimConfig()
    defaults = dict(SIZE_X=64, SIZE_Y=64, SIZE_Z=64, PADDING=3, GRAVITY=9.8, DT=1e-3, GAMMA=-7.0,
                    COMPRESSION=10.0, RESOLUTION=1.0, SUBSTEPS=20, SDF_RESOLUTION=64)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(c, k, v)
    _check(load_library().mpm_config_derive(ctypes.byref(c)))
    return c


def _parse(fn, arg: bytes, cap: int = 64):
    cfg = SimConfig()
    waters = (ConfigObject * cap)()
    objects = (ConfigObject * cap)()
    nw, no = ctypes.c_int(0), ctypes.c_int(0)
    _check(fn(arg, ctypes.byref(cfg), waters, ctypes.byref(nw), objects, ctypes.byref(no), cap))

    def conv(arr, n):
        return [dict(model_path=o.model_path.decode(), scale=o.scale, x=o.x, y=o.y, z=o.z)
                for o in list(arr)[: min(n, cap)]]

    return cfg, conv(waters, nw.value), conv(objects, no.value)


def parse_config_file(path) -> tuple[SimConfig, list, list]:
    """GameConfigData::setConfigDataFromFile + initDeviceParams: (config, waters, objects)."""
    return _parse(load_library().mpm_config_parse_file, str(path).encode())


def parse_config_string(text: str) -> tuple[SimConfig, list, list]:
    return _parse(load_library().mpm_config_parse_string, text.encode())


def config_get_int(text: str, key: str) -> int:
    v = ctypes.c_int64(0)
    _check(load_library().mpm_config_get_int(text.encode(), key.encode(), ctypes.byref(v)))
    return v.value


# ---- scene generators ----

def scene_fill(shape: str, lo, hi, seed: int) -> np.ndarray:
    """Jittered 8-per-cell particles as a (26, n) float32 array."""
    lib = load_library()
    code = {"box": 0, "sphere": 1}[shape]
    lo_a = _f32(lo, (3,))
    hi_a = _f32(np.resize(np.asarray(hi, dtype=np.float32), 3), (3,))
    cnt = ctypes.c_int64(0)
    _check(lib.mpm_scene_fill(code, _fptr(lo_a), _fptr(hi_a), seed, None, 0, ctypes.byref(cnt)))
    n = cnt.value
    out = np.zeros((PARTICLE_ATTRS, n), dtype=np.float32)
    cnt = ctypes.c_int64(0)
    _check(lib.mpm_scene_fill(code, _fptr(lo_a), _fptr(hi_a), seed, _fptr(out), n, ctypes.byref(cnt)))
    assert cnt.value == n
    return out


def _mesh(fn, *args) -> np.ndarray:
    n = ctypes.c_int(0)
    _check(fn(*args, None, ctypes.byref(n)))
    out = np.zeros((n.value, 9), dtype=np.float32)
    _check(fn(*args, _fptr(out), ctypes.byref(n)))
    return out


def mesh_uv_sphere(centre, radius: float, slices: int, stacks: int) -> np.ndarray:
    c = _f32(centre, (3,))
    return _mesh(load_library().mpm_mesh_uv_sphere, _fptr(c), ctypes.c_float(radius), slices, stacks)


def mesh_u_channel(lo, hi, wall: float, seg: float) -> np.ndarray:
    lo_a, hi_a = _f32(lo, (3,)), _f32(hi, (3,))
    return _mesh(load_library().mpm_mesh_u_channel, _fptr(lo_a), _fptr(hi_a), ctypes.c_float(wall), ctypes.c_float(seg))


def mesh_box(lo, hi) -> np.ndarray:
    lo_a, hi_a = _f32(lo, (3,)), _f32(hi, (3,))
    return _mesh(load_library().mpm_mesh_box, _fptr(lo_a), _fptr(hi_a))


# ---- scene front-end: OBJ meshes and config.txt ----

def obj_triangles(path) -> np.ndarray:
    """Triangles (T, 9) of an OBJ file (quads split along the shorter diagonal)."""
    lib = load_library()
    n = ctypes.c_int(0)
    _check(lib.mpm_obj_triangles(str(path).encode(), None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, 9), dtype=np.float32)
    _check(lib.mpm_obj_triangles(str(path).encode(), _fptr(out), n.value, ctypes.byref(n)))
    return out


def _points(fn, *args) -> np.ndarray:
    n = ctypes.c_int64(0)
    _check(fn(*args, None, 0, ctypes.byref(n)))
    out = np.zeros((3, n.value), dtype=np.float32)
    _check(fn(*args, _fptr(out), n.value, ctypes.byref(n)))
    return out


def voxelize_obj(path, resolution: float = 100.0) -> np.ndarray:
    """VoxelEngine::voxelize: (3, n) centres of the surface cells."""
    return _points(load_library().mpm_voxelize_obj, str(path).encode(), ctypes.c_float(resolution))


def obj_water(path, normalize_size: float, scale: float, displacement, lattice: int = 1) -> np.ndarray:
    """Simulation::Prepare_object: (3, n) water lattice points of a mesh."""
    d = _f32(displacement, (3,))
    return _points(load_library().mpm_obj_water, str(path).encode(), ctypes.c_float(normalize_size),
                   ctypes.c_float(scale), _fptr(d), int(lattice))


def obj_collider(path, normalize_size: float, scale: float, displacement) -> np.ndarray:
    """Simulation::Prepare_triangles: (T, 9) collider triangles of a mesh."""
    lib = load_library()
    d = _f32(displacement, (3,))
    n = ctypes.c_int(0)
    args = (str(path).encode(), ctypes.c_float(normalize_size), ctypes.c_float(scale), _fptr(d))
    _check(lib.mpm_obj_collider(*args, None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, 9), dtype=np.float32)
    _check(lib.mpm_obj_collider(*args, _fptr(out), n.value, ctypes.byref(n)))
    return out


def scene_from_config(path):
    """The reference's Simulation constructor without the GPU: (SimConfig,
    particles (26, n), triangles (T, 9)) from a config.txt and its models."""
    lib = load_library()
    cfg = SimConfig()
    n = ctypes.c_int64(0)
    nt = ctypes.c_int(0)
    p = str(path).encode()
    _check(lib.mpm_scene_from_config(p, ctypes.byref(cfg), None, 0, ctypes.byref(n), None, 0, ctypes.byref(nt)))
    parts = np.zeros((PARTICLE_ATTRS, n.value), dtype=np.float32)
    tris = np.zeros((nt.value, 9), dtype=np.float32)
    _check(lib.mpm_scene_from_config(p, ctypes.byref(cfg), _fptr(parts), n.value, ctypes.byref(n),
                                     _fptr(tris), nt.value, ctypes.byref(nt)))
    return cfg, parts, tris


# ---- engine ----

STAGES = ("p2g", "g2p", "g2p2g", "grid_update", "resort", "corrector")

CORRECTOR_FEATURES = 66
CORRECTOR_STRIDE = 72
CORRECTOR_PARAMS = 12803


class Engine:
    """The reference's Engine (engine.h:14-43) over libmpm.so.

    ``init(n, buffer)`` copies the caller's 26*n float buffer (the reference
    takes ownership of it instead); everything else keeps the reference's
    names and meaning.
    """

    def __init__(self, cfg: SimConfig, **options):
        self._lib = load_library()
        self.cfg = cfg
        self._nodes = cfg.nodes
        self._h = ctypes.c_void_p()
        _check(self._lib.mpm_engine_create(ctypes.byref(cfg), ctypes.byref(self._h)))
        for k, v in options.items():
            self.set_option(k, v)

    def set_dt_scale(self, scale: float) -> None:
        """Runs the following substeps at scale * DT (every scene)."""
        _check(self._lib.mpm_engine_set_dt_scale(self._h, float(scale)))

    def check(self, depth: float = 0.5) -> dict:
        """Fallback checks: particles deeper than `depth` in the collider,
        mean |det F - 1|, non-finite particles."""
        out = np.zeros(3, dtype=np.float64)
        _check(self._lib.mpm_engine_check(self._h, float(depth), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return {"penetrating": int(out[0]), "mean_abs_J_minus_1": float(out[1]), "nonfinite": int(out[2])}

    def set_corrector(self, corr) -> None:
        self._corr = corr  # keeps the corrector alive while attached
        _check(self._lib.mpm_engine_set_corrector(self._h, corr._h if corr is not None else None))

    def collect_pairs(self, ratio: int, cap: int):
        """(features (m, 72), targets (m, 3)) of one ratio*DT substep against
        `ratio` plain substeps; advances the engine by `ratio` plain substeps."""
        X = np.empty((cap, CORRECTOR_STRIDE), dtype=np.float32)
        T = np.empty((cap, 3), dtype=np.float32)
        cnt = ctypes.c_int(0)
        _check(self._lib.mpm_engine_collect_pairs(self._h, int(ratio), _fptr(X), _fptr(T), int(cap), ctypes.byref(cnt)))
        return X[:cnt.value].copy(), T[:cnt.value].copy()

    def debug_features(self, cap: int):
        """(nodes (m, 2) [x | y<<10 | z<<20, scene], features (m, 72)) of the
        active nodes of the current grid."""
        nodes = np.empty((cap, 2), dtype=np.int32)
        X = np.empty((cap, CORRECTOR_STRIDE), dtype=np.float32)
        cnt = ctypes.c_int(0)
        _check(self._lib.mpm_engine_debug_features(self._h, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                   _fptr(X), int(cap), ctypes.byref(cnt)))
        return nodes[:cnt.value].copy(), X[:cnt.value].copy()

    def snapshot(self) -> None:
        """Saves the particle state on the device (see restore)."""
        _check(self._lib.mpm_engine_snapshot(self._h))

    def restore(self) -> None:
        """Returns to the last snapshot, laid out for the current options."""
        _check(self._lib.mpm_engine_restore(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mpm_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, key: str, value: int) -> None:
        _check(self._lib.mpm_engine_set_option(self._h, key.upper().encode(), int(value)))

    def get_option(self, key: str) -> int:
        v = ctypes.c_int64(0)
        _check(self._lib.mpm_engine_get_option(self._h, key.upper().encode(), ctypes.byref(v)))
        return v.value

    # Engine::init (engine.cu:45-52)
    def init(self, n: int, buffer) -> None:
        buf = _f32(buffer).reshape(-1)
        if buf.size != PARTICLE_ATTRS * n:
            raise ValueError(f"buffer has {buf.size} floats, expected {PARTICLE_ATTRS * n}")
        _check(self._lib.mpm_engine_init_particles(self._h, n, _fptr(buf)))

    def load_state(self, state) -> None:
        """Resumes from a getParticles() state (F, v, C kept as they are)."""
        st = _f32(state).reshape(PARTICLE_ATTRS, -1)
        _check(self._lib.mpm_engine_load_state(self._h, st.shape[1], _fptr(np.ascontiguousarray(st))))

    # Engine::initSDF (engine.cu:105-114)
    def initSDF(self, triangles) -> None:
        t = _f32(triangles).reshape(-1, 9)
        _check(self._lib.mpm_engine_init_sdf_from_triangles(self._h, _fptr(t) if t.size else None, t.shape[0]))

    def set_sdf(self, sdf) -> None:
        s = _f32(sdf).reshape(-1)
        if s.size != self._nodes:
            raise ValueError("sdf must have one value per grid node")
        _check(self._lib.mpm_engine_set_sdf(self._h, _fptr(s)))

    def get_sdf(self) -> np.ndarray:
        out = np.empty(self._nodes, dtype=np.float32)
        _check(self._lib.mpm_engine_get_sdf(self._h, _fptr(out)))
        return out

    # Engine::step (engine.cu:54-70)
    def step(self) -> None:
        _check(self._lib.mpm_engine_step(self._h))

    def steps(self, k: int) -> None:
        _check(self._lib.mpm_engine_steps(self._h, int(k)))

    def synchronize(self) -> None:
        _check(self._lib.mpm_engine_synchronize(self._h))

    def getNumber(self) -> int:
        return int(self._lib.mpm_engine_get_number(self._h))

    def getParticles(self) -> np.ndarray:
        n = self.getNumber()
        out = np.empty((PARTICLE_ATTRS, n), dtype=np.float32)
        _check(self._lib.mpm_engine_get_particles(self._h, _fptr(out)))
        return out

    def getPositions(self, out=None) -> np.ndarray:
        """Positions (n, 3) in caller order; into `out` when given (a C-contiguous
        float32 (n, 3) array, e.g. pinned_array((n, 3)) for the full host-link
        rate)."""
        n = self.getNumber()
        if out is None:
            out = np.empty((n, 3), dtype=np.float32)
        elif out.shape != (n, 3) or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float32 array of shape (n, 3)")
        _check(self._lib.mpm_engine_get_positions(self._h, _fptr(out)))
        return out

    def getGrid(self) -> np.ndarray:
        out = np.empty((4, self._nodes), dtype=np.float32)
        _check(self._lib.mpm_engine_get_grid(self._h, _fptr(out)))
        return out

    # Engine::sortParticles (engine.cu:92-102): (sorted keys, permutation)
    def sortParticles(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.getNumber()
        keys = np.empty(n, dtype=np.int32)
        perm = np.empty(n, dtype=np.int32)
        _check(self._lib.mpm_engine_get_sort(
            self._h, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return keys, perm

    def stage_times(self) -> dict:
        out = np.zeros(2 * len(STAGES), dtype=np.float64)
        _check(self._lib.mpm_engine_get_stage_times(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(STAGES)))
        return {s: (float(out[2 * i]), int(out[2 * i + 1])) for i, s in enumerate(STAGES)}

    def counters(self) -> dict:
        out = np.zeros(5, dtype=np.int64)
        _check(self._lib.mpm_engine_get_counters(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 5))
        return {"chunks": int(out[0]), "tile_overflow_particles": int(out[1]), "active_blocks": int(out[2]),
                "claim_losers": int(out[3]), "rounds": int(out[4])}

    def debug_p2g(self) -> np.ndarray:
        out = np.empty((4, self._nodes), dtype=np.float32)
        _check(self._lib.mpm_engine_debug_p2g(self._h, _fptr(out)))
        return out

    def debug_grid_update(self, grid4) -> np.ndarray:
        g = _f32(grid4).reshape(4, -1)
        out = np.empty_like(g)
        _check(self._lib.mpm_engine_debug_grid_update(self._h, _fptr(g), _fptr(out)))
        return out

    def debug_g2p(self, grid4) -> None:
        g = _f32(grid4).reshape(4, -1)
        _check(self._lib.mpm_engine_debug_g2p(self._h, _fptr(g)))


class Batch(Engine):
    """Several independent scenes in one engine: every stage of a substep is
    one launch over all scenes (no reference counterpart).

    Particles of all scenes form one caller-order array, scene after scene;
    ``init(counts, buffer)`` takes the per-scene counts and the (26, sum)
    buffer.  The inherited read-back methods return that concatenation.
    """

    def __init__(self, cfgs, **options):
        self._lib = load_library()
        self.cfgs = list(cfgs)
        arr = (SimConfig * len(self.cfgs))(*self.cfgs)
        self.cfg = self.cfgs[0]
        self._nodes = sum(c.nodes for c in self.cfgs)
        self._h = ctypes.c_void_p()
        _check(self._lib.mpm_batch_create(arr, len(self.cfgs), ctypes.byref(self._h)))
        for k, v in options.items():
            self.set_option(k, v)

    @property
    def scenes(self) -> int:
        return len(self.cfgs)

    def init(self, counts, buffer) -> None:
        c = np.ascontiguousarray(counts, dtype=np.int32)
        if c.size != self.scenes:
            raise ValueError("one count per scene")
        buf = _f32(buffer).reshape(-1)
        if buf.size != PARTICLE_ATTRS * int(c.sum()):
            raise ValueError("buffer must hold 26 * sum(counts) floats")
        _check(self._lib.mpm_batch_init_particles(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                  _fptr(buf) if buf.size else None))

    def initSDF(self, scene: int, triangles) -> None:
        t = _f32(triangles).reshape(-1, 9)
        _check(self._lib.mpm_batch_init_sdf_from_triangles(self._h, int(scene), _fptr(t) if t.size else None,
                                                           t.shape[0]))

    def set_scene_sdf(self, scene: int, sdf) -> None:
        s = _f32(sdf).reshape(-1)
        if s.size != self.cfgs[scene].nodes:
            raise ValueError("sdf must have one value per node of the scene")
        _check(self._lib.mpm_batch_set_sdf(self._h, int(scene), _fptr(s)))

    def scene_info(self, scene: int) -> dict:
        out = np.zeros(4, dtype=np.int64)
        _check(self._lib.mpm_batch_scene_info(self._h, int(scene), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return {"first": int(out[0]), "count": int(out[1]), "node_offset": int(out[2]), "nodes": int(out[3])}

    def scene_grid(self, scene: int) -> np.ndarray:
        out = np.empty((4, self.cfgs[scene].nodes), dtype=np.float32)
        _check(self._lib.mpm_batch_get_grid(self._h, int(scene), _fptr(out)))
        return out


class Corrector:
    """The learned grid-velocity corrector's network on the device (no
    reference counterpart): parameters, input normalisation, forward pass,
    gradient and Adam training with the library's own kernels."""

    def __init__(self, params=None):
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        _check(self._lib.mpm_corrector_create(ctypes.byref(self._h)))
        if params is not None:
            self.set_params(params)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mpm_corrector_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params) -> None:
        p = _f32(params).reshape(-1)
        _check(self._lib.mpm_corrector_set_params(self._h, _fptr(p), p.size))

    def get_params(self) -> np.ndarray:
        p = np.empty(CORRECTOR_PARAMS, dtype=np.float32)
        _check(self._lib.mpm_corrector_get_params(self._h, _fptr(p), p.size))
        return p

    def set_norm(self, mean, inv_std, out_scale: float) -> None:
        m = _f32(mean, (CORRECTOR_FEATURES,))
        s = _f32(inv_std, (CORRECTOR_FEATURES,))
        _check(self._lib.mpm_corrector_set_norm(self._h, _fptr(m), _fptr(s), float(out_scale)))

    def set_clamp(self, clamp: float) -> None:
        _check(self._lib.mpm_corrector_set_clamp(self._h, float(clamp)))

    def set_deadzone(self, dead: float) -> None:
        """Outputs (before out_scale) of magnitude below `dead` give no correction."""
        _check(self._lib.mpm_corrector_set_deadzone(self._h, float(dead)))

    @staticmethod
    def _rows(X) -> np.ndarray:
        X = _f32(X)
        if X.ndim != 2 or X.shape[1] not in (CORRECTOR_FEATURES, CORRECTOR_STRIDE):
            raise ValueError("features must be (n, 66) or (n, 72)")
        if X.shape[1] == CORRECTOR_FEATURES:
            X = np.concatenate([X, np.zeros((X.shape[0], CORRECTOR_STRIDE - CORRECTOR_FEATURES), np.float32)], 1)
        return np.ascontiguousarray(X)

    def forward(self, X) -> np.ndarray:
        X = self._rows(X)
        Y = np.empty((X.shape[0], 3), dtype=np.float32)
        _check(self._lib.mpm_corrector_forward(self._h, _fptr(X), X.shape[0], _fptr(Y)))
        return Y

    def load_dataset(self, X, T) -> None:
        X = self._rows(X)
        T = _f32(T, (X.shape[0], 3))
        _check(self._lib.mpm_corrector_load_dataset(self._h, _fptr(X), _fptr(T), X.shape[0]))

    def gradient(self, row0: int, n: int):
        g = np.empty(CORRECTOR_PARAMS, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        _check(self._lib.mpm_corrector_gradient(self._h, int(row0), int(n), _fptr(g), ctypes.byref(loss)))
        return g, float(loss.value)

    def train(self, steps: int, batch: int, lr: float) -> np.ndarray:
        out = np.empty(steps, dtype=np.float32)
        _check(self._lib.mpm_corrector_train(self._h, int(steps), int(batch), float(lr), _fptr(out)))
        return out


class _PinnedBlock:
    """Owns one mpm_host_alloc block; freed when the last array using it goes."""

    def __init__(self, nbytes: int):
        self.lib = load_library()
        self.ptr = ctypes.c_void_p()
        _check(self.lib.mpm_host_alloc(int(nbytes), ctypes.byref(self.ptr)))

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self.lib.mpm_host_free(self.ptr)
            self.ptr = ctypes.c_void_p()


def pinned_array(shape, dtype=np.float32) -> np.ndarray:
    """A numpy array in page-locked host memory: read-backs into it (for
    example getPositions(out=...)) run at the full host-link rate.  The
    memory is freed with the last array that uses it."""
    dt = np.dtype(dtype)
    nbytes = max(1, int(np.prod(shape)) * dt.itemsize)
    block = _PinnedBlock(nbytes)
    buf = (ctypes.c_byte * nbytes).from_address(block.ptr.value)
    buf._owner = block  # the array's base keeps the block alive
    return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)


def probe_copy(nbytes: int, iters: int = 20, wide: bool = False) -> float:
    """Mean milliseconds of one device-to-device streaming copy of nbytes
    (wide: four 16-byte loads in flight per thread, nontemporal stores)."""
    ms = ctypes.c_double(0.0)
    fn = load_library().mpm_probe_copy_wide if wide else load_library().mpm_probe_copy
    _check(fn(int(nbytes), int(iters), ctypes.byref(ms)))
    return ms.value


def gpu_available() -> bool:
    """True when a HIP device is visible (torch is used only to ask)."""
    try:
        import torch

        return bool(torch.cuda.is_available())
    except Exception:
        return False


from . import scenes  # noqa: E402,F401
