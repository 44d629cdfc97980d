"""GPU tests of the reference-style program path: config.txt with WATER and
OBJECT meshes -> particles and collider -> engine (through the Python mirror
and through the headless driver), including checkpoint and resume."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def write_uv_sphere_obj(path, slices=16, stacks=12):
    """A unit UV sphere with quad faces (triangles at the poles)."""
    v = [(0.0, 1.0, 0.0)]
    for i in range(1, stacks):
        th = np.pi * i / stacks
        for j in range(slices):
            ph = 2 * np.pi * j / slices
            v.append((np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)))
    v.append((0.0, -1.0, 0.0))
    idx = lambda i, j: 2 + (i - 1) * slices + (j % slices)
    f = [(1, idx(1, j + 1), idx(1, j)) for j in range(slices)]
    for i in range(1, stacks - 1):
        for j in range(slices):
            f.append((idx(i, j), idx(i, j + 1), idx(i + 1, j + 1), idx(i + 1, j)))
    last = len(v)
    f += [(last, idx(stacks - 1, j), idx(stacks - 1, j + 1)) for j in range(slices)]
    path.write_text("".join(f"v {a:.6f} {b:.6f} {c:.6f}\n" for a, b, c in v) +
                    "".join("f " + " ".join(map(str, q)) + "\n" for q in f))


CONFIG = """SIZE_X=64
SIZE_Y=64
SIZE_Z=64

PADDING=2
GRAVITY=9.81
DT=0.002
GAMMA=-2.0
COMPRESSION=200.0
RESOLUTION=1.0
SUBSTEPS=10
SDF_RESOLUTION=64

WATER=sphere.obj;12.0;0.0;10.0;0.0
WATER=sphere.obj;8.0;14.0;14.0;10.0
OBJECT=bowl.obj;40.0;0.0;-20.0;0.0
"""


@pytest.fixture(scope="module")
def scene_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("scene")
    write_uv_sphere_obj(d / "sphere.obj")
    write_uv_sphere_obj(d / "bowl.obj", slices=24, stacks=16)
    (d / "config.txt").write_text(CONFIG)
    return d


def test_config_scene_one_substep_matches_oracle(gpu, oracle, scene_dir):
    cfg, parts, tris = gpu.scene_from_config(scene_dir / "config.txt")
    n = parts.shape[1]
    say(f"front-end scene: {n} particles, {len(tris)} collider triangles")
    assert n > 1000 and len(tris) > 100
    e = gpu.Engine(cfg)
    e.init(n, parts)
    e.initSDF(tris)
    sdf = e.get_sdf()
    ref_sdf = oracle.make_sdf(cfg, tris, nthreads=8)
    far = np.abs(ref_sdf) > 1e-3
    np.testing.assert_array_equal(np.sign(sdf[far]), np.sign(ref_sdf[far]))
    assert np.abs(np.abs(sdf) - np.abs(ref_sdf)).max() < 1e-4
    e.step()
    got = e.getParticles()
    ref = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, 1)
    assert float(np.abs(got[0:3] - ref[0:3]).max()) < 1e-4
    assert float(np.abs(got[3:6] - ref[3:6]).max()) < 1e-4
    e.close()


def run_headless(args, cwd):
    exe = Path(__import__("importlib").import_module("water-sim-cuda_amd").__file__).parent / "mpm_headless"
    r = subprocess.run([str(exe), *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]


def test_headless_checkpoint_resume(gpu, scene_dir, tmp_path):
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    full = run_headless(["config.txt", "--frames", "6", "--out", a, "--snapshots"], scene_dir)
    first = run_headless(["config.txt", "--frames", "3", "--out", b, "--checkpoint"], scene_dir)
    rest = run_headless(["config.txt", "--frames", "3", "--out", b, "--resume", b / "state_000003.bin"], scene_dir)
    say(f"headless: {full[-1]}")
    assert [r["frame"] for r in full] == list(range(1, 7))
    assert [r["frame"] for r in rest] == [4, 5, 6]
    n = full[0]["particles"]
    for r in full:
        assert r["nonfinite"] == 0 and abs(r["grid_mass"] - n) < 1e-3 * n
    # resumed run = uninterrupted run (float-atomic order aside)
    for x, y in zip(full[3:], rest):
        assert abs(x["mean_y"] - y["mean_y"]) < 1e-3 and abs(x["max_y"] - y["max_y"]) < 1e-2
        assert abs(x["time"] - y["time"]) < 1e-9
    # snapshot format: header + n x 3 positions, matching the stats
    raw = (a / "frame_000006.pos").read_bytes()
    assert raw[:4] == b"MPMP"
    pos = np.frombuffer(raw[24:], dtype=np.float32).reshape(-1, 3)
    assert pos.shape[0] == n
    assert abs(float(pos[:, 1].mean()) - full[-1]["mean_y"]) < 1e-4
    assert len((b / "stats.jsonl").read_text().splitlines()) == 6
