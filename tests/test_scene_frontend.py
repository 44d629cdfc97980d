"""CPU tests of the scene front-end (OBJ -> water particles and collider
triangles, config.txt -> buffers; csrc/voxel.cpp).

Pins: hand-worked quad splits and index forms; the voxeliser's separating-
axis test against an exact brute force (clip the triangle to the cell,
Sutherland-Hodgman in float64: non-empty <=> overlap) on random meshes; the
closed-form surface-cell count of an axis-aligned cube (with the reference's
max-plane quirk); the closed-form
lattice shell of a normalised, dilated cube; the affine map and culling of
collider triangles; the packed buffer of a config with two WATER entries.
"""
import numpy as np
import pytest

CUBE_QUADS = """# unit cube, quads
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
vn 0 0 1
vt 0 0
f 1/1/1 4/1/1 3/1/1 2/1/1
f 5//1 6//1 7//1 8//1
f 1 2 6 5
f 2 3 7 6
f 3 4 8 7
f -4 -8 -5 -1
"""


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_obj_quads_split_along_shorter_diagonal(mpm, tmp_path):
    # quad 0-1-2-3 whose 1-3 diagonal is shorter than 0-2
    p = write(tmp_path, "q.obj", "v 0 0 0\nv 2 -1 0\nv 4 0 0\nv 2 1 0\nf 1 2 3 4\n"
                                 "v 0 0 5\nv 1 0 5\nv 1 1 5\nv 0 1 5\nf 5/1 6/2 7/3 8/4\n")
    t = mpm.obj_triangles(p).reshape(-1, 3, 3)
    assert t.shape == (4, 3, 3)
    # |v2 - v0|^2 = 16 > |v3 - v1|^2 = 4: triangles (0,1,3), (1,2,3)
    np.testing.assert_array_equal(t[0], [[0, 0, 0], [2, -1, 0], [2, 1, 0]])
    np.testing.assert_array_equal(t[1], [[2, -1, 0], [4, 0, 0], [2, 1, 0]])
    # unit square: diagonals tie -> (0,1,3), (1,2,3) as well (the test is strict <)
    np.testing.assert_array_equal(t[2], [[0, 0, 5], [1, 0, 5], [0, 1, 5]])
    # |v2 - v0|^2 = 12.5 < |v3 - v1|^2 = 18: the other split, (0,1,2), (0,2,3)
    p2 = write(tmp_path, "r.obj", "v 0 0 0\nv 3 0 0\nv 2.5 2.5 0\nv 0 3 0\nf 1 2 3 4\n")
    t2 = mpm.obj_triangles(p2).reshape(-1, 3, 3)
    np.testing.assert_array_equal(t2[0], [[0, 0, 0], [3, 0, 0], [2.5, 2.5, 0]])
    np.testing.assert_array_equal(t2[1], [[0, 0, 0], [2.5, 2.5, 0], [0, 3, 0]])


def test_obj_index_forms_and_negative_indices(mpm, tmp_path):
    t = mpm.obj_triangles(write(tmp_path, "c.obj", CUBE_QUADS)).reshape(-1, 3, 3)
    assert t.shape == (12, 3, 3)
    # the last face uses relative indices -4 -8 -5 -1 = vertices 5 1 4 8 (x = 0 face)
    np.testing.assert_array_equal(t[-2:, :, 0], 0.0)
    with pytest.raises(mpm.MPMError):
        mpm.obj_triangles(write(tmp_path, "bad.obj", "v 0 0 0\nf 1 2 3\n"))
    with pytest.raises(mpm.MPMError):
        mpm.obj_triangles(tmp_path / "missing.obj")


def clip_overlap(tri, lo, hi):
    """Exact triangle/box overlap in float64: clip the triangle polygon by the
    six half-spaces of the box; the box and triangle overlap iff something
    (a point, a segment or an area) remains."""
    poly = [np.asarray(v, dtype=np.float64) for v in tri]
    for ax in range(3):
        for side, bound in ((1.0, lo[ax]), (-1.0, hi[ax])):
            out = []
            n = len(poly)
            for i in range(n):
                a, b = poly[i], poly[(i + 1) % n]
                da, db = side * (a[ax] - bound), side * (b[ax] - bound)
                if da >= 0:
                    out.append(a)
                if (da >= 0) != (db >= 0):
                    t = da / (da - db)
                    out.append(a + t * (b - a))
            poly = out
            if not poly:
                return False
    return True


def test_voxelize_matches_exact_clipping(mpm, tmp_path):
    rng = np.random.default_rng(11)
    for trial in range(3):
        pts = rng.uniform(-1.0, 1.0, size=(12, 3)).astype(np.float32)
        faces = rng.integers(0, 12, size=(10, 3))
        faces = faces[(faces[:, 0] != faces[:, 1]) & (faces[:, 1] != faces[:, 2]) & (faces[:, 0] != faces[:, 2])]
        text = "".join(f"v {p[0]!r} {p[1]!r} {p[2]!r}\n" for p in pts.tolist())
        text += "".join(f"f {a + 1} {b + 1} {c + 1}\n" for a, b, c in faces)
        path = write(tmp_path, f"soup{trial}.obj", text)
        got = mpm.voxelize_obj(path, 12.0)
        tris = pts[faces]
        # the bounding box covers every vertex of the file, used or not
        lo = pts.min(0)
        cell = np.float32((pts.max(0) - lo).max() / np.float32(12.0))
        g = np.ceil((pts.max(0) - lo) / cell).astype(int)
        ref = set()
        for t in tris:
            i0 = np.maximum(0, np.floor((t.min(0) - lo) / cell).astype(int))
            i1 = np.minimum(g - 1, np.ceil((t.max(0) - lo) / cell).astype(int))
            for z in range(i0[2], i1[2] + 1):
                for y in range(i0[1], i1[1] + 1):
                    for x in range(i0[0], i1[0] + 1):
                        bmin = lo + np.array([x, y, z], np.float32) * cell
                        if clip_overlap(t, bmin.astype(np.float64), (bmin + cell).astype(np.float64)):
                            ref.add((x, y, z))
        cells = {tuple(int(round(float(c))) for c in ((p - lo) / cell - 0.5)) for p in got.T}
        assert cells == ref, (trial, len(cells ^ ref))


def test_voxelize_cube_closed_form(mpm, tmp_path):
    """An axis-aligned unit cube at resolution n: the faces on the bounding
    box's min planes give every cell with a zero index.  The faces on the max
    planes give none: their cell range starts at floor((max - min) / cell),
    which in FP32 is n, one past the last cell (VoxelEngine.cpp:164-169), so
    the reference never tests them.  Count: n^3 - (n - 1)^3."""
    path = write(tmp_path, "cube.obj", CUBE_QUADS)
    for res in (10, 37, 100):
        got = mpm.voxelize_obj(path, float(res))
        assert got.shape[1] == res ** 3 - (res - 1) ** 3, res
        cell = np.float32(1.0) / np.float32(res)
        idx = np.round(got / cell - 0.5).astype(int)
        assert (idx.min(0) == 0).all() and (idx == 0).any(0).all()
        assert ((idx == 0).any(0)).all()


def test_water_lattice_shell_closed_form(mpm, tmp_path):
    """The cube's 29,701 surface voxels (100^3 - 99^3: the three min faces),
    normalised to a 64 grid, scaled to 10 cells, displaced by (20, 20, 20):
    the faces land on the lattice planes x = 47, y = 47, z = 47 over
    [47, 57]^2; dilated by 2 they cover the points of [45, 59]^3 with some
    coordinate <= 49: 15^3 - 10^3 = 2375."""
    path = write(tmp_path, "cube.obj", CUBE_QUADS)
    w = mpm.obj_water(path, 64.0, 10.0, (20.0, 20.0, 20.0), 1)
    assert w.shape[1] == 15 ** 3 - 10 ** 3
    assert (w.min(1) == 45).all() and (w.max(1) == 59).all()
    assert (w.min(0) <= 49).all()
    # (x, y, z) lexicographic order of the unique lattice points
    keys = (w[0].astype(np.int64) * 4096 + w[1].astype(np.int64)) * 4096 + w[2].astype(np.int64)
    assert (np.diff(keys) > 0).all()
    # clamped to [0, 64]: a displacement past the far wall snaps the x = 67
    # face to 64, and the dilation reaches 62
    w2 = mpm.obj_water(path, 64.0, 10.0, (40.0, 0.0, 0.0), 1)
    assert w2[0].max() == 64 and w2[0].min() == 62
    # RESOLUTION below 1 (read as integer 0 by the reference) is rejected
    with pytest.raises(mpm.MPMError):
        mpm.obj_water(path, 64.0, 10.0, (0.0, 0.0, 0.0), 0)


def test_collider_affine_map_and_culling(mpm, tmp_path):
    path = write(tmp_path, "cube.obj", CUBE_QUADS)
    t = mpm.obj_collider(path, 64.0, 10.0, (20.0, 20.0, 20.0)).reshape(-1, 3, 3)
    assert t.shape == (12, 3, 3)
    for ax in range(3):
        assert set(np.unique(t[:, :, ax]).tolist()) == {47.0, 57.0}
    # y in {57, 67}: the two triangles of the y = 67 face lie wholly outside [0, 64]
    assert mpm.obj_collider(path, 64.0, 10.0, (20.0, 30.0, 20.0)).shape[0] == 10
    # all vertices beyond x = 64: every triangle is dropped
    assert mpm.obj_collider(path, 64.0, 10.0, (100.0, 0.0, 0.0)).shape[0] == 0
    # one face straddles the wall: triangles with a vertex inside stay
    t3 = mpm.obj_collider(path, 64.0, 10.0, (32.0, 0.0, 0.0)).reshape(-1, 3, 3)
    assert 0 < t3.shape[0] < 12 and (t3[:, :, 0].min(1) <= 64.0).all()


def test_scene_from_config(mpm, tmp_path):
    write(tmp_path, "cube.obj", CUBE_QUADS)
    cfg_text = ("SIZE_X=64\nSIZE_Y=64\nSIZE_Z=64\nPADDING=3\nGRAVITY=9.81\nDT=0.001\nGAMMA=-2.0\n"
                "COMPRESSION=200.0\nRESOLUTION=1.0\nSUBSTEPS=10\nSDF_RESOLUTION=64\n"
                "WATER=cube.obj;10.0;20.0;30.0;20.0\nWATER=cube.obj;6.0;0.0;0.0;0.0\n"
                "OBJECT=cube.obj;20.0;0.0;-20.0;0.0\n")
    cfgp = write(tmp_path, "config.txt", cfg_text)
    cfg, parts, tris = mpm.scene_from_config(cfgp)
    a = mpm.obj_water(tmp_path / "cube.obj", 64.0, 10.0, (20.0, 30.0, 20.0), 1)
    b = mpm.obj_water(tmp_path / "cube.obj", 64.0, 6.0, (0.0, 0.0, 0.0), 1)
    n = a.shape[1] + b.shape[1]
    assert parts.shape == (26, n)
    np.testing.assert_array_equal(parts[0:3], np.concatenate([a, b], 1))
    assert not parts[3:24].any() and (parts[24:26] == 1.0).all()
    np.testing.assert_array_equal(tris, mpm.obj_collider(tmp_path / "cube.obj", 64.0, 20.0, (0.0, -20.0, 0.0)))
    assert cfg.SIZE_X == 64 and abs(cfg.DT - 0.001) < 1e-9


def test_reference_scene_fixture(mpm):
    """tests/golden/reference_scene.npz holds the reference program's own
    scene (config.txt:1-18 with models/sphere.obj twice and models/u.obj):
    one particle per lattice node of the two dilated sphere shells, merged
    without de-duplication (Simulation::MergeVoxelData, Simulation.cpp:122-160),
    VOL = MASS = 1 and zero elsewhere (VoxelPipeline.cpp:9-30), and the U's
    quads split into triangles inside the grid.  Where the reference tree is
    present, the front-end rebuilds it bit for bit."""
    from pathlib import Path

    golden = Path(__file__).resolve().parent / "golden"
    d = np.load(golden / "reference_scene.npz")
    cfg, parts, tris = d["cfg"], d["particles"], d["triangles"]
    assert list(d["cfg_fields"][:3]) == ["SIZE_X", "SIZE_Y", "SIZE_Z"] and list(cfg[:4]) == [128, 128, 128, 2]
    n = parts.shape[1]
    assert parts.shape == (26, n) and n > 5000
    assert np.array_equal(parts[0:3], np.round(parts[0:3]))  # RESOLUTION 1: integer lattice
    assert (parts[24] == 1).all() and (parts[25] == 1).all() and not parts[3:24].any()
    # two copies of one sphere shell, the second shifted by (10, 0, 0): equal halves
    half = n // 2
    assert n == 2 * half
    a, b = parts[0:3, :half], parts[0:3, half:]
    np.testing.assert_array_equal(np.sort(b[0] - 10.0), np.sort(a[0]))
    assert tris.shape[1] == 9 and 0 < tris.shape[0] <= 576  # u.obj: 288 quads -> 576 triangles, culled
    ref = Path("/root/reference/config.txt")
    if ref.exists():
        sys_path = str(golden)
        import sys

        sys.path.insert(0, sys_path)
        try:
            import make_reference_scene as mk
        finally:
            sys.path.remove(sys_path)
        vals, p2, t2 = mk.build(ref)
        np.testing.assert_array_equal(vals, cfg)
        np.testing.assert_array_equal(p2, parts)
        np.testing.assert_array_equal(t2, tris)
