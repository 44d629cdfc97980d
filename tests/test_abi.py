"""Host-side checks of libmpm.so that need no GPU: the library loads and
exports every entry point that include/mpm.h declares, the config parser
follows the reference grammar (game_configdata.cpp:22-90), and the scene
generators are deterministic, closed and outward-wound."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from helpers import signed_volume

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "mpm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(mpm):
    lib = ctypes.CDLL(str(mpm.library_path()))
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol(mpm):
    lib = mpm.load_library()
    for n in declared_functions():
        assert getattr(lib, n).argtypes is not None, n


REF_STYLE = """SIZE_X=128
SIZE_Y = 96
SIZE_Z=64

PADDING=2
GRAVITY=9.81
DT=0.01
GAMMA=-2.0
COMPRESSION=200.0
RESOLUTION=1.0
SUBSTEPS=10
SDF_RESOLUTION=256
RESORT_EVERY=10
this line has no equals sign

WATER=models/sphere.obj;15.0;0.0;20.0;25.0
WATER = models/sphere.obj ; 15.0 ; 10.0 ; 20.0 ; 25.0
OBJECT=models/u.obj;110.0;0.0;-10.0;0.0
"""


def test_config_parse_reference_grammar(mpm):
    cfg, waters, objects = mpm.parse_config_string(REF_STYLE)
    assert (cfg.SIZE_X, cfg.SIZE_Y, cfg.SIZE_Z, cfg.PADDING) == (128, 96, 64, 2)
    assert cfg.GRAVITY == pytest.approx(9.81) and cfg.DT == pytest.approx(0.01)
    assert cfg.GAMMA == pytest.approx(-2.0) and cfg.COMPRESSION == pytest.approx(200.0)
    assert (cfg.SUBSTEPS, cfg.SDF_RESOLUTION) == (10, 256)
    # derived keys, Simulation.cpp:70-72
    assert cfg.GRID_NUMBER == 128 * 96 * 64
    assert cfg.GRID_SIZE == 128 * 96 * 64 * 16
    assert cfg.GRID_BLOCKS == (128 * 96 * 64 + 255) // 256
    assert len(waters) == 2 and len(objects) == 1
    assert waters[1]["model_path"] == "models/sphere.obj"  # spaces removed
    assert waters[1]["x"] == pytest.approx(10.0)
    assert objects[0]["scale"] == pytest.approx(110.0) and objects[0]["y"] == pytest.approx(-10.0)
    assert mpm.config_get_int(REF_STYLE, "RESORT_EVERY") == 10


def test_config_parse_file(mpm, tmp_path):
    p = tmp_path / "config.txt"
    p.write_text(REF_STYLE)
    cfg, waters, _ = mpm.parse_config_file(p)
    assert cfg.SIZE_Y == 96 and len(waters) == 2


def test_config_missing_key_is_an_error(mpm):
    text = REF_STYLE.replace("DT=0.01\n", "")
    with pytest.raises(mpm.MPMError, match="DT"):
        mpm.parse_config_string(text)


def test_config_malformed_object_is_an_error(mpm):
    with pytest.raises(mpm.MPMError, match="WATER"):
        mpm.parse_config_string(REF_STYLE + "WATER=a.obj;1.0;2.0\n")


def test_config_missing_file(mpm, tmp_path):
    with pytest.raises(mpm.MPMError):
        mpm.parse_config_file(tmp_path / "nope.txt")


def test_scene_fill_box_counts_and_bounds(mpm):
    p = mpm.scene_fill("box", (2.0, 3.0, 4.0), (5.0, 7.0, 6.0), 11)
    assert p.shape == (26, 8 * 3 * 4 * 2)
    assert (p[0] >= 2.0).all() and (p[0] < 5.0).all()
    assert (p[1] >= 3.0).all() and (p[1] < 7.0).all()
    np.testing.assert_array_equal(p[3:24], 0.0)
    np.testing.assert_array_equal(p[24:26], 1.0)
    # stratified: exactly one particle per 0.5-cell sub-cell
    sub = np.floor(p[0:3] * 2).astype(int)
    assert len({tuple(c) for c in sub.T}) == p.shape[1]


def test_scene_fill_deterministic(mpm):
    a = mpm.scene_fill("sphere", (32.0, 40.0, 32.0), (12.0,), 1)
    b = mpm.scene_fill("sphere", (32.0, 40.0, 32.0), (12.0,), 1)
    c = mpm.scene_fill("sphere", (32.0, 40.0, 32.0), (12.0,), 2)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a[0:3], c[0:3][:, : a.shape[1]]) if c.shape == a.shape else True
    # 8 per cell inside r=12: about 8 * 4/3 pi 12^3 = 57,906
    assert abs(a.shape[1] - 57906) < 600
    r = np.linalg.norm(a[0:3] - np.array([[32.0], [40.0], [32.0]], dtype=np.float32), axis=0)
    assert r.max() < 12.0


def test_uv_sphere_mesh(mpm):
    t = mpm.mesh_uv_sphere((0.0, 0.0, 0.0), 20.0, 32, 33)
    assert t.shape == (2048, 9)
    v = signed_volume(t)
    assert 0.95 * (4 / 3) * np.pi * 20 ** 3 < v < (4 / 3) * np.pi * 20 ** 3


def test_box_mesh(mpm):
    t = mpm.mesh_box((1.0, 2.0, 3.0), (2.5, 4.0, 7.0))
    assert t.shape == (12, 9)
    assert signed_volume(t) == pytest.approx(1.5 * 2.0 * 4.0, rel=1e-6)


def test_u_channel_mesh_volume(mpm):
    lo, hi, wall = np.array([6.3, 20.3, 6.3]), np.array([249.7, 90.3, 249.7]), 4.0
    t = mpm.mesh_u_channel(lo, hi, wall, 6.25)
    assert 9000 < t.shape[0] < 11000
    L = hi[0] - lo[0]
    H = hi[1] - lo[1]
    W = hi[2] - lo[2]
    # U cross-section = outer rectangle minus the slot above the floor
    area = W * H - (W - 2 * wall) * (H - wall)
    assert signed_volume(t) == pytest.approx(L * area, rel=1e-5)


def test_engine_without_gpu_raises(mpm):
    if mpm.gpu_available():
        pytest.skip("GPU present")
    with pytest.raises(mpm.MPMError):
        mpm.Engine(mpm.make_config())
