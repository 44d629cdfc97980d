"""GPU tests of the accelerated makeSDF (bounding-volume hierarchy): the same
field as the brute-force makeSDF (kernels.cu:282-300), which the engine also
runs when SDF_BVH = 0.

Distances must agree bit for bit (the hierarchy only skips triangles that a
box test excludes; the minimum over the rest is the same float).  Signs must
agree wherever the brute-force distance is non-zero: the approximate winding
sum decides only where it is clearly away from 1/2, and every other node
takes the exact sum in input order.  Includes the full 256^3 channel mesh
(16.7 M nodes, ~10 k triangles), where the hierarchy must also be faster.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def both_sdfs(gpu, cfg, tris):
    out = []
    for bvh in (0, 1):
        e = gpu.Engine(cfg, SDF_BVH=bvh)
        t = time.perf_counter()
        e.initSDF(tris)
        dt = time.perf_counter() - t
        out.append((e.get_sdf(), dt, e.get_option("SDF_EXACT_NODES")))
        e.close()
    return out


def check_same(brute, bvh):
    np.testing.assert_array_equal(np.abs(bvh), np.abs(brute))
    nz = brute != 0.0
    np.testing.assert_array_equal(np.signbit(bvh[nz]), np.signbit(brute[nz]))


def perturbed_sphere(gpu, rng):
    """A closed, star-shaped, non-convex mesh: a UV sphere with radial noise."""
    t = gpu.mesh_uv_sphere((0.0, 0.0, 0.0), 1.0, 40, 41).reshape(-1, 3, 3).astype(np.float64)
    v, inv = np.unique(t.reshape(-1, 3).round(6), axis=0, return_inverse=True)
    r = 1.0 + 0.25 * np.sin(5 * v[:, 0]) * np.cos(4 * v[:, 1]) + 0.05 * rng.standard_normal(len(v))
    v = v * (9.0 * r)[:, None] + np.array([21.0, 19.5, 20.5])
    return v[inv.reshape(-1)].reshape(-1, 9).astype(np.float32)


@pytest.mark.parametrize("mesh", ["sphere", "box", "channel", "perturbed", "open_bowl", "soup"])
def test_bvh_sdf_equals_brute_force(gpu, mesh):
    say(f"bvh makeSDF: {mesh}")
    cfg = gpu.make_config(SIZE_X=44, SIZE_Y=40, SIZE_Z=42)
    rng = np.random.default_rng(5)
    if mesh == "sphere":
        tris = gpu.mesh_uv_sphere((21.3, 19.8, 20.6), 12.0, 48, 49)
    elif mesh == "box":
        tris = gpu.mesh_box((4.3, 3.7, 5.2), (35.6, 19.4, 27.1))
    elif mesh == "channel":
        tris = gpu.mesh_u_channel((2.3, 3.3, 2.3), (41.7, 25.3, 39.7), 3.0, 4.0)
    elif mesh == "perturbed":
        tris = perturbed_sphere(gpu, rng)
    elif mesh == "open_bowl":
        # the lower half of a sphere: an open surface, winding numbers between
        # 0 and 1 all around it
        t = gpu.mesh_uv_sphere((21.0, 22.0, 20.0), 13.0, 36, 36).reshape(-1, 3, 3)
        tris = t[t[:, :, 1].max(1) <= 22.0 + 1e-4].reshape(-1, 9)
    else:
        tris = rng.uniform(2.0, 38.0, size=(300, 9)).astype(np.float32)
    (brute, tb, _), (bvh, tv, exact) = both_sdfs(gpu, cfg, tris)
    say(f"  {len(tris)} triangles: brute {tb * 1e3:.1f} ms, bvh {tv * 1e3:.1f} ms, exact-sign nodes {exact}")
    check_same(brute, bvh)


def test_bvh_sdf_channel256_full_size(gpu):
    cfg = gpu.make_config(SIZE_X=256, SIZE_Y=256, SIZE_Z=256)
    tris = gpu.mesh_u_channel((6.3, 20.3, 6.3), (249.7, 90.3, 249.7), 4.0, 6.25)
    (brute, tb, _), (bvh, tv, exact) = both_sdfs(gpu, cfg, tris)
    say(f"channel256 makeSDF ({len(tris)} triangles): brute {tb * 1e3:.1f} ms, bvh {tv * 1e3:.1f} ms, "
        f"exact-sign nodes {exact}")
    check_same(brute, bvh)
    assert tv < tb / 5
