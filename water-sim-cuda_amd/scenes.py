"""Benchmark and test scenes (seeded, generated; no reference counterpart).

Each scene returns a dict with ``cfg`` (SimConfig), ``particles`` (26 x n
float32, caller order), the collider as either ``triangles`` (T x 9, built
into the SDF by makeSDF) or ``sdf`` (G floats, analytic), ``resort_every``,
``frames`` and a short ``description``.

Parameter conventions.  The reference subtracts GRAVITY*DT from v_y
(kernels.cu:155) and computes pressure as COMPRESSION*(J^GAMMA - 1) clamped
at zero with J = det F (kernels.cu:26-29).  Physical water under these
formulas needs GRAVITY > 0 (downwards) and GAMMA < 0 (pressure under
compression): ``physical=True`` uses GRAVITY = 9.8 and GAMMA = -7 (Tait's
exponent 7 on density rho/rho0 = 1/J).  ``physical=False`` keeps the literal
values GRAVITY = -9.8, GAMMA = 7 of the 64^3 sphere workload; parity does not
depend on the choice because both sides receive the same values.
"""
from __future__ import annotations

import numpy as np


def _pkg():
    import importlib

    return importlib.import_module("water-sim-cuda_amd")


def _params(physical: bool) -> dict:
    if physical:
        return dict(GRAVITY=9.8, GAMMA=-7.0, COMPRESSION=10.0)
    return dict(GRAVITY=-9.8, GAMMA=7.0, COMPRESSION=10.0)


def floor_sdf(cfg, top: float) -> np.ndarray:
    """Analytic axis-aligned box floor: node SDF = y - top (negative below)."""
    sz, sy, sx = cfg.grid_shape()
    y = np.arange(sy, dtype=np.float32)
    return np.broadcast_to((y - np.float32(top))[None, :, None], (sz, sy, sx)).astype(np.float32).reshape(-1)


def sphere64(physical: bool = False, seed: int = 1, radius: float = 12.0) -> dict:
    """64^3, PADDING 3, water sphere r=12 at (32,40,32), 8 ppc, DT 1e-3,
    20 substeps x 100 frames, analytic floor at y = 4."""
    m = _pkg()
    cfg = m.make_config(SIZE_X=64, SIZE_Y=64, SIZE_Z=64, PADDING=3, DT=1e-3, RESOLUTION=1.0, SUBSTEPS=20,
                        SDF_RESOLUTION=64, **_params(physical))
    parts = m.scene_fill("sphere", (32.0, 40.0, 32.0), (radius,), seed)
    return dict(name="sphere64", cfg=cfg, particles=parts, sdf=floor_sdf(cfg, 4.0), triangles=None,
                resort_every=1, frames=100,
                description="64^3 grid, water sphere r=12 at (32,40,32), 8 ppc, box floor SDF")


def dambreak128(physical: bool = True, seed: int = 2) -> dict:
    """128^3 dam break: 48x96x48-cell block against the x=4 wall (N = 1,769,472),
    UV-sphere collider (32 slices x 33 stacks = 2,048 triangles, r = 20),
    DT 5e-4, 20 substeps x 200 frames, re-sort every 4 substeps; the whole
    run is autotuned (its splash phase wants smaller chunks than the fresh
    block)."""
    m = _pkg()
    cfg = m.make_config(SIZE_X=128, SIZE_Y=128, SIZE_Z=128, PADDING=3, DT=5e-4, RESOLUTION=1.0, SUBSTEPS=20,
                        SDF_RESOLUTION=128, **_params(physical))
    parts = m.scene_fill("box", (4.0, 4.0, 40.0), (52.0, 100.0, 88.0), seed)
    tris = m.mesh_uv_sphere((88.3, 24.3, 64.3), 20.0, 32, 33)
    return dict(name="dambreak128", cfg=cfg, particles=parts, sdf=None, triangles=tris, resort_every=4,
                frames=200, tune=True, description="128^3 dam break 48x96x48 cells, UV-sphere r=20 (2048 tris)")


def channel256(physical: bool = True, seed: int = 3) -> dict:
    """256^3: water boxes 96x128x96 and 64^3 cells (N = 11,534,336) above a
    U-channel (wall 4 cells, ~10k triangles), DT 2.5e-4, 40 substeps x 100
    frames.  This scene measures the HBM roofline."""
    m = _pkg()
    cfg = m.make_config(SIZE_X=256, SIZE_Y=256, SIZE_Z=256, PADDING=3, DT=2.5e-4, RESOLUTION=1.0, SUBSTEPS=40,
                        SDF_RESOLUTION=256, **_params(physical))
    a = m.scene_fill("box", (16.0, 104.0, 16.0), (112.0, 232.0, 112.0), seed)
    b = m.scene_fill("box", (150.0, 120.0, 150.0), (214.0, 184.0, 214.0), seed + 1000)
    parts = np.concatenate([a, b], axis=1)
    tris = m.mesh_u_channel((6.3, 20.3, 6.3), (249.7, 90.3, 249.7), 4.0, 6.25)
    return dict(name="channel256", cfg=cfg, particles=parts, sdf=None, triangles=tris, resort_every=5,
                frames=100, description="256^3, boxes 96x128x96 + 64^3 cells into a U-channel mesh")


def small_box(n_cells=(6, 6, 6), grid=32, physical: bool = True, seed: int = 7, dt: float = 1e-3) -> dict:
    """A small jittered block for quick tests (8 * prod(n_cells) particles)."""
    m = _pkg()
    cfg = m.make_config(SIZE_X=grid, SIZE_Y=grid, SIZE_Z=grid, PADDING=3, DT=dt, RESOLUTION=1.0, SUBSTEPS=20,
                        SDF_RESOLUTION=grid, **_params(physical))
    lo = np.array([grid / 2 - n_cells[0] / 2, 8.0, grid / 2 - n_cells[2] / 2], dtype=np.float32)
    hi = lo + np.array(n_cells, dtype=np.float32)
    parts = m.scene_fill("box", lo, hi, seed)
    return dict(name="small_box", cfg=cfg, particles=parts, sdf=floor_sdf(cfg, 4.0), triangles=None,
                resort_every=1, frames=10, description=f"{grid}^3 block {n_cells}")


def sphere_sdf(cfg, centre, radius: float) -> np.ndarray:
    """Analytic solid-sphere collider: node SDF = |x - centre| - radius."""
    sz, sy, sx = cfg.grid_shape()
    z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    d = np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2) - radius
    return d.astype(np.float32).reshape(-1)


def sample_sdf(cfg, sdf: np.ndarray, pos: np.ndarray) -> np.ndarray:
    """Trilinear node-SDF values at (3, n) positions, cells clamped into the
    grid as getSDF does (common.cuh:136-167); float64."""
    sx, sy, sz = int(cfg.SIZE_X), int(cfg.SIZE_Y), int(cfg.SIZE_Z)
    p = np.asarray(pos, dtype=np.float64)
    i0 = np.floor(p).astype(np.int64)
    t = p - i0
    dims = np.array([sx, sy, sz])[:, None]
    lo = np.clip(i0, 0, dims - 1)
    hi = np.clip(i0 + 1, 0, dims - 1)
    s = np.asarray(sdf, dtype=np.float64).reshape(-1)
    out = np.zeros(p.shape[1])
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                x = hi[0] if cx else lo[0]
                y = hi[1] if cy else lo[1]
                z = hi[2] if cz else lo[2]
                w = ((t[0] if cx else 1 - t[0]) * (t[1] if cy else 1 - t[1]) * (t[2] if cz else 1 - t[2]))
                out += w * s[x + sx * (y + sy * z)]
    return out


def batch_scenes(count: int = 64, seed: int = 4, n_min: int = 5000, n_max: int = 400000,
                 edges=(32, 48, 64, 96), physical: bool = True) -> dict:
    """Independent scenes for the batched engine.

    Per scene (seeded): grid edge uniform in ``edges``; a target particle
    count log-uniform in [n_min, n_max], capped at 0.3 * 8 * edge^3 so that it
    fits the grid (a 400k-particle scene does not fit a 32^3 grid at 8 particles
    per cell); a water sphere or cube of that volume at 8 ppc placed at random
    in the upper half; a random collider, either a box floor (y below a random
    height) or a solid sphere below the water, as an analytic node SDF.
    Particles that would start inside the collider (node SDF interpolated at
    their position below zero) are left out, so no water starts in a solid.
    DT 1e-3, 20 substeps per frame, PADDING 3.

    Returns ``cfgs``, ``counts``, ``particles`` (26 x sum(counts), scene after
    scene), ``sdfs`` (one array per scene), ``resort_every`` (the fastest
    measured period with the autotuner), ``chunk`` (the starting CHUNK before
    tuning: the scenes are small and splash, so smaller work items balance
    better late in a run) and ``description``.
    """
    m = _pkg()
    rng = np.random.default_rng(seed)
    cfgs, parts, sdfs, counts = [], [], [], []
    for s in range(count):
        e = int(rng.choice(edges))
        cfg = m.make_config(SIZE_X=e, SIZE_Y=e, SIZE_Z=e, PADDING=3, DT=1e-3, RESOLUTION=1.0, SUBSTEPS=20,
                            SDF_RESOLUTION=e, **_params(physical))
        target = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        target = min(target, 0.3 * 8 * e ** 3)
        cells = target / 8.0
        margin = 3 + 2
        if rng.random() < 0.5:
            r = (3.0 * cells / (4.0 * np.pi)) ** (1.0 / 3.0)
            r = min(r, e / 2 - margin - 0.5)
            lo_c, hi_c = margin + r, e - margin - r
            c = [rng.uniform(lo_c, hi_c), rng.uniform(max(lo_c, e / 2 - 0.5 * r), hi_c), rng.uniform(lo_c, hi_c)]
            p = m.scene_fill("sphere", c, (r,), seed * 1000 + s)
        else:
            a = min(cells ** (1.0 / 3.0), e - 2 * margin - 1)
            lo_c, hi_c = margin, e - margin - a
            lo = np.array([rng.uniform(lo_c, hi_c), rng.uniform(max(lo_c, min(e / 2 - a / 2, hi_c)), hi_c),
                           rng.uniform(lo_c, hi_c)], dtype=np.float32)
            lo = np.floor(lo)
            p = m.scene_fill("box", lo, lo + np.float32(np.floor(a)), seed * 1000 + s)
        if rng.random() < 0.5:
            sdf = floor_sdf(cfg, float(rng.uniform(4.0, e / 4)))
        else:
            rc = float(rng.uniform(e / 8, e / 5))
            sdf = sphere_sdf(cfg, (rng.uniform(e / 3, 2 * e / 3), 3.0 + 0.5 * rc, rng.uniform(e / 3, 2 * e / 3)), rc)
        p = np.ascontiguousarray(p[:, sample_sdf(cfg, sdf, p[0:3]) >= 0.0])
        cfgs.append(cfg)
        parts.append(p)
        sdfs.append(sdf)
        counts.append(p.shape[1])
    return dict(name=f"batch{count}", cfgs=cfgs, counts=np.array(counts, dtype=np.int32),
                particles=np.concatenate(parts, axis=1), sdfs=sdfs, resort_every=4, chunk=256, frames=100, tune=True,
                description=f"{count} scenes, edges {edges}, {n_min}-{n_max} particles log-uniform (capped at "
                            f"0.3*8*edge^3), sphere/cube water, floor or sphere collider, DT 1e-3")


SCENES = {"sphere64": sphere64, "dambreak128": dambreak128, "channel256": channel256}
