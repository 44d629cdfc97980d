"""Shared helpers for the tests (independent restatements used as checks)."""
from __future__ import annotations

import numpy as np


def config(oracle, **kw):
    c = oracle.Config()
    base = dict(SIZE_X=32, SIZE_Y=32, SIZE_Z=32, PADDING=3, GRAVITY=9.8, DT=1e-3, GAMMA=-7.0, COMPRESSION=10.0,
                RESOLUTION=1.0, SUBSTEPS=20, SDF_RESOLUTION=32)
    base.update(kw)
    for k, v in base.items():
        setattr(c, k, v)
    g = c.SIZE_X * c.SIZE_Y * c.SIZE_Z
    c.GRID_NUMBER = g
    c.GRID_SIZE = g * 16
    c.GRID_BLOCKS = (g + 255) // 256
    return c


def morton_bruteforce(x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Bit-by-bit interleave (bit i of x -> 3i, of y -> 3i+1, of z -> 3i+2),
    10 bits per axis; independent of the multiply-and-mask spreading."""
    x = x.astype(np.uint64); y = y.astype(np.uint64); z = z.astype(np.uint64)
    out = np.zeros_like(x)
    for i in range(10):
        out |= ((x >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i)
        out |= ((y >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i + 1)
        out |= ((z >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i + 2)
    return out.astype(np.uint32)


def keys_numpy(pos: np.ndarray) -> np.ndarray:
    """setKeys (kernels.cu:361-371) for positions inside [0, 1024)^3."""
    c = np.floor(np.maximum(pos.astype(np.float64), 0.0)).astype(np.int64)
    assert (c < 1024).all()
    return morton_bruteforce(c[0], c[1], c[2]).astype(np.int32)


def stable_perm(keys: np.ndarray) -> np.ndarray:
    return np.argsort(keys.astype(np.int64), kind="stable").astype(np.int32)


def single_particle(n_attrs=26, pos=(10.3, 12.7, 11.1), vel=(0.5, -1.0, 2.0), C=None, F=None, dtype=np.float64):
    p = np.zeros((n_attrs, 1), dtype=dtype)
    p[0:3, 0] = pos
    p[3:6, 0] = vel
    if C is not None:
        p[6:15, 0] = np.asarray(C, dtype=dtype).reshape(9)
    p[15:24, 0] = np.eye(3).reshape(9) if F is None else np.asarray(F, dtype=dtype).reshape(9)
    p[24, 0] = 1.0
    p[25, 0] = 1.0
    return p


def node_coords(cfg):
    sx, sy, sz = int(cfg.SIZE_X), int(cfg.SIZE_Y), int(cfg.SIZE_Z)
    z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    return np.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)]).astype(np.float64)


def signed_volume(tris: np.ndarray) -> float:
    """Divergence theorem: sum of signed tetra volumes (origin, v0, v1, v2)."""
    t = tris.astype(np.float64).reshape(-1, 3, 3)
    return float(np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2])).sum() / 6.0)


def host_threads() -> int:
    """Host CPUs this process may use (affinity set, capped by a cgroup CPU
    quota), for the OpenMP oracle."""
    import os
    from pathlib import Path

    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(quota) // int(period)))
    except (OSError, ValueError):
        pass
    return max(1, n)


def parity_log(record: dict) -> None:
    """Appends a JSON line to $MPM_PARITY_LOG when it is set (the per-config
    error table of the full-size runs)."""
    import json
    import os

    path = os.environ.get("MPM_PARITY_LOG")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(record) + "\n")
