"""Writes tests/golden/reference_scene.npz: the reference program's own scene.

    python tests/golden/make_reference_scene.py   (build container only)

The reference's application reads config.txt (config.txt:1-18: a 128^3 grid,
two WATER spheres from models/sphere.obj, one OBJECT from models/u.obj) and
its Simulation constructor turns it into the particle buffer handed to
Engine::init and the triangles handed to Engine::initSDF (Simulation.cpp:4-38,
VoxelEngine.cpp, VoxelPipeline.cpp:5-33).  This script runs the same front-end
of this library (mpm_scene_from_config) on those files where they lie under
/root/reference and stores its output as data: the config values, the 26 x n
particle buffer and the collider triangles.  The GPU tests run that scene
where the reference tree is absent; tests/test_scene_frontend.py checks the
fixture against a fresh front-end run wherever the reference tree is present.
"""
import importlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

REFERENCE_CONFIG = Path("/root/reference/config.txt")
OUT = Path(__file__).resolve().parent / "reference_scene.npz"
CFG_FIELDS = ("SIZE_X", "SIZE_Y", "SIZE_Z", "PADDING", "GRAVITY", "DT", "GAMMA", "COMPRESSION", "RESOLUTION",
              "SUBSTEPS", "SDF_RESOLUTION")


def build(config_path: Path = REFERENCE_CONFIG):
    m = importlib.import_module("water-sim-cuda_amd")
    cfg, parts, tris = m.scene_from_config(config_path)
    vals = np.array([float(getattr(cfg, k)) for k in CFG_FIELDS], dtype=np.float64)
    return vals, np.ascontiguousarray(parts, dtype=np.float32), np.ascontiguousarray(tris, dtype=np.float32)


def main() -> None:
    if not REFERENCE_CONFIG.exists():
        sys.exit(f"{REFERENCE_CONFIG} not found: the fixture is written in the build container only")
    vals, parts, tris = build()
    np.savez_compressed(OUT, cfg_fields=np.array(CFG_FIELDS), cfg=vals, particles=parts, triangles=tris)
    print(f"wrote {OUT}: {parts.shape[1]} particles, {tris.shape[0]} triangles, {OUT.stat().st_size} bytes")


if __name__ == "__main__":
    main()
