"""The reference program's own scene on the GPU.

tests/golden/reference_scene.npz is the scene the reference's Simulation
constructor builds from its config.txt (config.txt:1-18: 128^3 grid, DT 0.01,
GAMMA -2, COMPRESSION 200, two models/sphere.obj waters, a models/u.obj
collider), as written by tests/golden/make_reference_scene.py.  Here it runs
through the engine: the collider SDF against the oracle's brute-force
makeSDF, one substep against the FP32 and FP64 oracles, and ten frames
(100 substeps, Simulation::run) through the headless driver with the
invariants a correct step keeps and the statistics of the oracle's own ten
frames.
"""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from helpers import host_threads, parity_log

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_scene.npz"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def max_err(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) if a.size else 0.0


@pytest.fixture(scope="module")
def scene(gpu):
    d = np.load(GOLDEN)
    kw = {k: float(v) for k, v in zip(d["cfg_fields"], d["cfg"])}
    for k in ("SIZE_X", "SIZE_Y", "SIZE_Z", "PADDING", "SUBSTEPS", "SDF_RESOLUTION"):
        kw[k] = int(kw[k])
    cfg = gpu.make_config(**kw)
    return cfg, d["particles"], d["triangles"]


def test_reference_scene_sdf_and_one_substep(gpu, oracle, scene):
    cfg, parts, tris = scene
    n = parts.shape[1]
    say(f"reference scene: {n} particles, {len(tris)} triangles, grid {cfg.SIZE_X}^3, DT {cfg.DT}")
    e = gpu.Engine(cfg)
    e.init(n, parts)
    e.initSDF(tris)
    sdf = e.get_sdf()
    ref_sdf = oracle.make_sdf(cfg, tris, nthreads=host_threads())
    far = np.abs(ref_sdf) > 1e-3
    np.testing.assert_array_equal(np.sign(sdf[far]), np.sign(ref_sdf[far]))
    assert np.abs(np.abs(sdf) - np.abs(ref_sdf)).max() < 1e-4
    e.step()
    got = e.getParticles()
    e.close()
    ref32 = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, 1)
    ref64 = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float64), sdf.astype(np.float64), 1)
    gaps = {"gpu_vs_fp32": {"pos": max_err(got[0:3], ref32[0:3]), "vel": max_err(got[3:6], ref32[3:6])},
            "gpu_vs_fp64": {"pos": max_err(got[0:3], ref64[0:3]), "vel": max_err(got[3:6], ref64[3:6])},
            "fp32_vs_fp64": {"pos": max_err(ref32[0:3], ref64[0:3]), "vel": max_err(ref32[3:6], ref64[3:6])}}
    say(f"reference scene one substep: {gaps}")
    parity_log({"workload": "reference config.txt", "particles": n, **gaps})
    assert gaps["gpu_vs_fp32"]["pos"] < 1e-4 and gaps["gpu_vs_fp32"]["vel"] < 1e-4
    assert gaps["gpu_vs_fp64"]["pos"] < 1e-4 and gaps["gpu_vs_fp64"]["vel"] < 1e-4


def write_scene(path: Path, cfg, parts, tris) -> None:
    """The headless driver's scene file (MPMC, see csrc/headless.cpp)."""
    import ctypes

    n, nt = parts.shape[1], tris.shape[0]
    with open(path, "wb") as f:
        f.write(b"MPMC" + np.array([1, n, nt], dtype=np.uint32).tobytes())
        f.write(bytes(ctypes.string_at(ctypes.addressof(cfg), ctypes.sizeof(cfg))))
        f.write(np.ascontiguousarray(parts, dtype=np.float32).tobytes())
        f.write(np.ascontiguousarray(tris, dtype=np.float32).tobytes())


def test_reference_scene_ten_frames_headless(gpu, oracle, scene, tmp_path):
    cfg, parts, tris = scene
    n = parts.shape[1]
    path = tmp_path / "reference.mpmc"
    write_scene(path, cfg, parts, tris)
    exe = Path(gpu.__file__).parent / "mpm_headless"
    r = subprocess.run([str(exe), "--scene", str(path), "--frames", "10", "--out", str(tmp_path), "--snapshots"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    say(f"headless reference scene: {rows[-1]}")
    assert [row["frame"] for row in rows] == list(range(1, 11))
    for row in rows:
        assert row["particles"] == n and row["nonfinite"] == 0 and row["penetrating"] == 0
        assert abs(row["grid_mass"] - n) < 1e-3 * n
    assert abs(rows[-1]["time"] - 10 * int(cfg.SUBSTEPS) * float(cfg.DT)) < 1e-6
    raw = (tmp_path / "frame_000010.pos").read_bytes()
    pos = np.frombuffer(raw[24:], dtype=np.float32).reshape(-1, 3).T
    # the oracle's own ten frames (Simulation::run's loop, engine.cu:54-70 per substep)
    e = gpu.Engine(cfg)
    e.init(n, parts)
    e.initSDF(tris)
    sdf = e.get_sdf()
    e.close()
    ref = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, 10 * int(cfg.SUBSTEPS),
                       nthreads=host_threads())
    gy, oy = pos[1].astype(np.float64), ref[1].astype(np.float64)
    d = {"mean_y": abs(gy.mean() - oy.mean()), "std_y": abs(gy.std() - oy.std()),
         "p05_y": abs(np.percentile(gy, 5) - np.percentile(oy, 5)),
         "p95_y": abs(np.percentile(gy, 95) - np.percentile(oy, 95)),
         "max_pos": max_err(pos, ref[0:3])}
    say(f"reference scene after 10 frames: gpu mean_y {gy.mean():.4f} oracle {oy.mean():.4f}; gaps {d}")
    parity_log({"workload": "reference config.txt", "frames": 10, "gaps": d})
    assert abs(oy.mean() - parts[1].mean()) > 1.0  # the water fell
    assert d["mean_y"] < 0.25 and d["std_y"] < 0.25 and d["p05_y"] < 0.5 and d["p95_y"] < 0.5
