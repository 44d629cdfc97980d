"""Pins the CPU oracle (a restatement of the reference step) with closed forms,
invariants and brute force.  Each check is chosen so that a plausible
mistake -- a dropped term, a transposed C, a wrong sign or index -- fails.
Runs on the CPU."""
import math

import numpy as np
import pytest

from helpers import config, keys_numpy, morton_bruteforce, node_coords, signed_volume, single_particle, stable_perm


# ---------------- Morton keys and sort (integer, bit-exact) ----------------

def test_morton_closed_forms(oracle):
    L = oracle.lib()
    assert L.oracle_morton(1, 0, 0) == 1
    assert L.oracle_morton(0, 1, 0) == 2
    assert L.oracle_morton(0, 0, 1) == 4
    assert L.oracle_morton(1023, 1023, 1023) == (1 << 30) - 1
    assert L.oracle_morton(2, 0, 0) == 8
    assert L.oracle_expand_bits(0b1111111111) == 0x09249249


def test_morton_bruteforce(oracle):
    rng = np.random.default_rng(0)
    xyz = rng.integers(0, 1024, size=(3, 5000))
    got = np.array([oracle.lib().oracle_morton(int(a), int(b), int(c)) for a, b, c in xyz.T], dtype=np.uint32)
    np.testing.assert_array_equal(got, morton_bruteforce(xyz[0], xyz[1], xyz[2]))


def test_keys_from_positions(oracle):
    rng = np.random.default_rng(1)
    pos = rng.uniform(-3.0, 300.0, size=(3, 4000)).astype(np.float32)
    pos[:, :5] = [[0.0, 0.999, 1.0, -0.5, 255.99]] * 3
    np.testing.assert_array_equal(oracle.morton_keys(pos), keys_numpy(pos))


def test_sort_is_stable_ascending(oracle):
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 50, size=3000).astype(np.int32)  # many ties
    keys[::7] = -5  # signed order
    perm = oracle.sort_perm(keys)
    ref = sorted(range(keys.size), key=lambda i: (int(keys[i]), i))
    np.testing.assert_array_equal(perm, np.array(ref, dtype=np.int32))


# ---------------- quadratic B-spline (common.cuh:205-210) ----------------

def test_spline_values(oracle):
    assert oracle.spline(0.0) == pytest.approx(0.75)
    assert oracle.spline(1.0) == pytest.approx(0.125)
    assert oracle.spline(-1.0) == pytest.approx(0.125)
    assert oracle.spline(0.5) == pytest.approx(0.5)
    assert oracle.spline(1.5) == 0.0
    assert oracle.spline(2.0) == 0.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spline_moments(oracle, dtype):
    """Partition of unity, zero first moment, second moment 1/4 on the stencil
    base = trunc(x - 0.5), for x >= 0.5."""
    rng = np.random.default_rng(3)
    for x in rng.uniform(0.5, 40.0, size=300):
        x = float(np.float32(x))
        b = int(x - 0.5)
        w = [oracle.spline(x - (b + i), dtype) for i in range(3)]
        d = [(b + i) - x for i in range(3)]
        tol = 1e-6 if dtype == np.float32 else 1e-12
        assert sum(w) == pytest.approx(1.0, abs=tol)
        assert sum(wi * di for wi, di in zip(w, d)) == pytest.approx(0.0, abs=tol * 4)
        assert sum(wi * di * di for wi, di in zip(w, d)) == pytest.approx(0.25, abs=tol * 4)


# ---------------- P2G (kernels.cu:6-142) ----------------

def _p2g_moments(cfg, grid, xp):
    X = node_coords(cfg)
    m = grid[0]
    mom = grid[1:4]
    d = X - np.asarray(xp, dtype=np.float64)[:, None]
    return m.sum(), mom.sum(axis=1), mom @ d.T, (m * X).sum(axis=1)


def test_p2g_single_particle_moments(oracle):
    """Mass, linear momentum and the affine moment of one particle:
    sum m_i = m, sum p_i = m v, sum p_i (x_i - x_p)^T = m C D with D = I/4,
    sum m_i x_i = m x_p.  A transposed or mis-scaled C changes the third."""
    cfg = config(oracle, GAMMA=-7.0, COMPRESSION=10.0)
    C = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [0.5, 0.9, -0.6]])
    xp = (10.3, 12.7, 11.1)
    v = (0.5, -1.0, 2.0)
    p = single_particle(pos=xp, vel=v, C=C)  # F = I -> J = 1 -> pressure 0
    g = oracle.p2g(cfg, p)
    msum, psum, aff, mx = _p2g_moments(cfg, g, xp)
    assert msum == pytest.approx(1.0, abs=1e-12)
    np.testing.assert_allclose(psum, v, atol=1e-12)
    np.testing.assert_allclose(aff, C / 4.0, atol=1e-12)
    np.testing.assert_allclose(mx, xp, atol=1e-12)


@pytest.mark.parametrize("s", [0.98, 0.95, 0.7])
def test_p2g_pressure_term(oracle, s):
    """With F = s*I the reference adds 4*DT*V*J*p/m * I to C, p =
    COMPRESSION*(J^GAMMA - 1) with J = det F clamped to [0.9, 1.1]."""
    cfg = config(oracle, GAMMA=-7.0, COMPRESSION=10.0, DT=1e-3)
    xp = (9.6, 8.2, 10.9)
    p = single_particle(pos=xp, vel=(0.0, 0.0, 0.0), F=np.eye(3) * s)
    g = oracle.p2g(cfg, p)
    _, psum, aff, _ = _p2g_moments(cfg, g, xp)
    # the reference clamps with the float literals 0.9f and 1.1f
    J = min(max(s ** 3, float(np.float32(0.9))), float(np.float32(1.1)))
    pressure = max(10.0 * (J ** -7.0 - 1.0), 0.0)
    multi = 4.0 * float(np.float32(1e-3)) * 1.0 * J * pressure
    np.testing.assert_allclose(aff, np.eye(3) * multi / 4.0, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(psum, 0.0, atol=1e-12)


def test_p2g_mass_conservation_many(oracle):
    """Interior particles: total grid mass = N (partition of unity)."""
    cfg = config(oracle)
    rng = np.random.default_rng(4)
    n = 2000
    p = np.zeros((26, n), dtype=np.float64)
    p[0:3] = rng.uniform(6.0, 26.0, size=(3, n))
    p[3:6] = rng.normal(size=(3, n))
    p = oracle.init_state(cfg, p, np.float64)
    g = oracle.p2g(cfg, p)
    assert g[0].sum() == pytest.approx(n, rel=1e-12)
    np.testing.assert_allclose(g[1:4].sum(axis=1), p[3:6].sum(axis=1), rtol=1e-9, atol=1e-9)


def test_p2g_boundary_nodes_skipped(oracle):
    """Nodes outside the grid receive nothing: a particle at x = 0.7 (base 0)
    keeps only in-bounds weights."""
    cfg = config(oracle)
    p = single_particle(pos=(0.7, 10.5, 10.5))
    g = oracle.p2g(cfg, p)
    w = [oracle.spline(0.7 - i, np.float64) for i in range(3)]
    assert g[0].sum() == pytest.approx(sum(w), abs=1e-12)


# ---------------- grid update (kernels.cu:144-196) ----------------

def _grid_with(cfg, entries):
    G = cfg.SIZE_X * cfg.SIZE_Y * cfg.SIZE_Z
    g = np.zeros((4, G), dtype=np.float64)
    for (x, y, z), m, mom in entries:
        i = z * cfg.SIZE_X * cfg.SIZE_Y + y * cfg.SIZE_X + x
        g[0, i] = m
        g[1:4, i] = mom
    return g


def _idx(cfg, x, y, z):
    return z * cfg.SIZE_X * cfg.SIZE_Y + y * cfg.SIZE_X + x


def test_grid_update_closed_forms(oracle):
    cfg = config(oracle, GRAVITY=9.8, DT=1e-3, PADDING=3)
    G = cfg.SIZE_X * cfg.SIZE_Y * cfg.SIZE_Z
    sdf = np.full(G, 100.0)
    g = _grid_with(cfg, [((10, 10, 10), 2.0, (1.0, 4.0, -2.0)),   # free node
                         ((1, 10, 10), 1.0, (-3.0, 0.5, 0.0)),    # x < PAD, moving -x
                         ((30, 10, 10), 1.0, (3.0, 0.5, 0.0)),    # x > X-1-PAD, moving +x
                         ((12, 12, 12), 1e-10, (5.0, 5.0, 5.0)),  # below mass threshold
                         ((5, 2, 5), 1.0, (0.0, -1.0, 0.0))])     # y < PAD moving down
    out = oracle.grid_update(cfg, g, sdf)
    gdt = float(np.float32(9.8)) * float(np.float32(1e-3))
    np.testing.assert_allclose(out[1:4, _idx(cfg, 10, 10, 10)], (0.5, 2.0 - gdt, -1.0), atol=1e-12)
    np.testing.assert_allclose(out[1:4, _idx(cfg, 1, 10, 10)], (0.0, 0.5 - gdt, 0.0), atol=1e-12)
    np.testing.assert_allclose(out[1:4, _idx(cfg, 30, 10, 10)], (0.0, 0.5 - gdt, 0.0), atol=1e-12)
    np.testing.assert_allclose(out[1:4, _idx(cfg, 12, 12, 12)], 0.0)
    np.testing.assert_allclose(out[1:4, _idx(cfg, 5, 2, 5)], 0.0, atol=1e-12)
    assert out[0, _idx(cfg, 10, 10, 10)] == 2.0  # mass kept


def test_grid_update_slip_against_plane(oracle):
    """Planar SDF y - 8.5: nodes below move with their normal component
    removed when approaching, untouched when separating (free slip)."""
    cfg = config(oracle, GRAVITY=0.0)
    X = node_coords(cfg)
    sdf = X[1] - 8.5
    g = _grid_with(cfg, [((10, 8, 10), 1.0, (1.0, -2.0, 0.5)), ((14, 8, 14), 1.0, (1.0, 2.0, 0.5)),
                         ((16, 12, 16), 1.0, (1.0, -2.0, 0.5))])
    out = oracle.grid_update(cfg, g, sdf)
    np.testing.assert_allclose(out[1:4, _idx(cfg, 10, 8, 10)], (1.0, 0.0, 0.5), atol=1e-7)
    np.testing.assert_allclose(out[1:4, _idx(cfg, 14, 8, 14)], (1.0, 2.0, 0.5), atol=1e-12)
    np.testing.assert_allclose(out[1:4, _idx(cfg, 16, 12, 16)], (1.0, -2.0, 0.5), atol=1e-12)


# ---------------- G2P (kernels.cu:199-280) ----------------

def _affine_grid(cfg, v0, A):
    X = node_coords(cfg)
    g = np.zeros((4, X.shape[1]))
    g[0] = 1.0
    g[1:4] = np.asarray(v0)[:, None] + np.asarray(A) @ X
    return g


def test_g2p_affine_round_trip(oracle):
    """An affine grid velocity v0 + A x is reproduced exactly: v_p = v0 + A x_p
    and C_p = 4 A D = A (a transposed C would give A^T); F' = (I + DT A) F."""
    cfg = config(oracle, DT=1e-3)
    G = cfg.SIZE_X * cfg.SIZE_Y * cfg.SIZE_Z
    A = np.array([[0.2, -0.5, 0.1], [0.3, 0.05, -0.7], [-0.4, 0.6, 0.15]])
    v0 = np.array([1.0, -2.0, 0.5])
    grid = _affine_grid(cfg, v0, A)
    rng = np.random.default_rng(5)
    n = 50
    p = np.zeros((26, n))
    p[0:3] = rng.uniform(6.0, 26.0, size=(3, n))
    F0 = np.array([[1.01, 0.02, 0.0], [-0.01, 0.99, 0.03], [0.0, 0.01, 1.0]])
    p[15:24] = F0.reshape(9, 1)
    p[24:26] = 1.0
    out = oracle.g2p(cfg, p, grid, np.full(G, 1e15))
    dt = float(np.float32(1e-3))
    vel = v0[:, None] + A @ p[0:3]
    np.testing.assert_allclose(out[3:6], vel, atol=1e-11)
    np.testing.assert_allclose(out[6:15], np.repeat(A.reshape(9, 1), n, axis=1), atol=1e-11)
    np.testing.assert_allclose(out[0:3], p[0:3] + vel * dt, atol=1e-11)
    Fn = (np.eye(3) + dt * A) @ F0
    np.testing.assert_allclose(out[15:24], np.repeat(Fn.reshape(9, 1), n, axis=1), atol=1e-12)


def test_g2p_volume_floor(oracle):
    """det F' < 0.3 rescales F' isotropically so det F' = 0.3."""
    cfg = config(oracle)
    G = cfg.SIZE_X * cfg.SIZE_Y * cfg.SIZE_Z
    grid = _affine_grid(cfg, (0, 0, 0), np.zeros((3, 3)))
    p = single_particle(pos=(12.2, 13.4, 14.6), F=np.diag([0.5, 0.6, 0.7]))
    out = oracle.g2p(cfg, p, grid, np.full(G, 1e15))
    Fn = out[15:24, 0].reshape(3, 3)
    floor = float(np.float32(0.3))  # the literal 0.3f
    assert np.linalg.det(Fn) == pytest.approx(floor, rel=1e-12)
    scale = (floor / (0.5 * 0.6 * 0.7)) ** (1.0 / 3.0)
    np.testing.assert_allclose(Fn, np.diag([0.5, 0.6, 0.7]) * scale, atol=1e-12)


def test_g2p_collider_projection(oracle):
    """A particle advected below the plane y = 8.5 is pushed back onto it
    and loses its approaching normal velocity."""
    cfg = config(oracle, DT=0.1)
    X = node_coords(cfg)
    sdf = X[1] - 8.5
    grid = _affine_grid(cfg, (1.0, -3.0, 0.0), np.zeros((3, 3)))
    p = single_particle(pos=(12.2, 8.6, 14.6))
    out = oracle.g2p(cfg, p, grid, sdf)
    # calculateNormal: central difference with eps 0.1f, divided by len + 1e-6f
    eps = float(np.float32(0.1))
    ny = 2 * eps / (2 * eps + float(np.float32(1e-6)))
    y_adv = 8.6 + (-3.0) * float(np.float32(0.1))
    assert out[1, 0] == pytest.approx(y_adv - ny * (y_adv - 8.5), abs=1e-9)
    np.testing.assert_allclose(out[3:6, 0], (1.0, -3.0 * (1.0 - ny * ny), 0.0), atol=1e-9)


# ---------------- whole substep (engine.cu:54-70) ----------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lone_particle_free_fall(oracle, dtype):
    cfg = config(oracle, GRAVITY=9.8, DT=1e-3)
    G = cfg.SIZE_X * cfg.SIZE_Y * cfg.SIZE_Z
    p = single_particle(pos=(15.3, 16.7, 14.2), vel=(0.3, 1.0, -0.2), dtype=dtype)
    state = oracle.init_state(cfg, p, dtype)
    out = oracle.steps(cfg, state, np.full(G, 1e15, dtype=dtype), 3)
    gdt = float(np.float32(9.8)) * float(np.float32(1e-3))
    dt = float(np.float32(1e-3))
    v = np.array([0.3, 1.0, -0.2])
    x = np.array([15.3, 16.7, 14.2])
    for _ in range(3):
        v = v - np.array([0.0, gdt, 0.0])
        x = x + v * dt
    tol = 1e-9 if dtype == np.float64 else 2e-5
    np.testing.assert_allclose(out[3:6, 0], v, atol=tol)
    np.testing.assert_allclose(out[0:3, 0], x, atol=tol)
    np.testing.assert_allclose(out[6:15, 0], 0.0, atol=tol)


def _small_scene(oracle, mpm_free=True):
    cfg = config(oracle, SIZE_X=24, SIZE_Y=24, SIZE_Z=24, GRAVITY=9.8, GAMMA=-7.0, COMPRESSION=10.0, DT=2e-3)
    rng = np.random.default_rng(6)
    n = 3000
    p = np.zeros((26, n), dtype=np.float32)
    p[0:3] = rng.uniform(6.0, 14.0, size=(3, n))
    p[1] += 2.0
    p[24:26] = 1.0
    X = node_coords(cfg)
    sdf = (X[1] - 4.0).astype(np.float32)
    return cfg, p, sdf


def test_fp32_matches_fp64(oracle):
    cfg, p, sdf = _small_scene(oracle)
    s32 = oracle.init_state(cfg, p, np.float32)
    s64 = oracle.init_state(cfg, p, np.float64)
    o32 = oracle.steps(cfg, s32, sdf, 10)
    o64 = oracle.steps(cfg, s64, sdf.astype(np.float64), 10)
    assert np.abs(o32[0:3] - o64[0:3]).max() < 1e-4
    assert np.abs(o32[3:6] - o64[3:6]).max() < 1e-3


def test_openmp_matches_serial(oracle):
    cfg, p, sdf = _small_scene(oracle)
    s = oracle.init_state(cfg, p, np.float32)
    a = oracle.steps(cfg, s, sdf, 5, nthreads=1)
    b = oracle.steps(cfg, s, sdf, 5, nthreads=4)
    assert np.abs(a[0:3] - b[0:3]).max() < 1e-5
    assert np.abs(a[15:24] - b[15:24]).max() < 1e-5


def test_mass_and_momentum_without_gravity(oracle):
    """No gravity, no collider contact, interior particles: the grid after
    P2G carries total mass N and total momentum sum(m v) every substep."""
    cfg, p, sdf = _small_scene(oracle)
    cfg.GRAVITY = 0.0
    sdf = np.full_like(sdf, 1e15)
    s = oracle.init_state(cfg, p, np.float64)
    out = oracle.steps(cfg, s, sdf.astype(np.float64), 3)
    g = oracle.p2g(cfg, out)
    assert g[0].sum() == pytest.approx(p.shape[1], rel=1e-12)
    np.testing.assert_allclose(g[1:4].sum(axis=1), out[3:6].sum(axis=1), rtol=1e-9, atol=1e-9)


# ---------------- mesh SDF (kernels.cu:282-300) ----------------

def test_solid_angle_octant(oracle):
    tri = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float32)
    assert oracle.solid_angle((0, 0, 0), tri) == pytest.approx(math.pi / 2, rel=1e-6)
    tri2 = np.array([1, 0, 0, 0, 0, 1, 0, 1, 0], dtype=np.float32)  # reversed winding
    assert oracle.solid_angle((0, 0, 0), tri2) == pytest.approx(-math.pi / 2, rel=1e-6)


def test_point_triangle_distance_bruteforce(oracle):
    rng = np.random.default_rng(8)
    s = np.linspace(0.0, 1.0, 301)
    u, v = np.meshgrid(s, s)
    mask = u + v <= 1.0
    u, v = u[mask], v[mask]
    for _ in range(40):
        tri = rng.normal(size=(3, 3)) * 3.0
        P = rng.normal(size=3) * 4.0
        pts = tri[0] + u[:, None] * (tri[1] - tri[0]) + v[:, None] * (tri[2] - tri[0])
        brute = ((pts - P) ** 2).sum(axis=1).min()
        d2 = oracle.point_tri_dist2(P, tri.reshape(9), np.float64)
        assert d2 <= brute + 1e-9
        h = np.linalg.norm(tri[1] - tri[0]) + np.linalg.norm(tri[2] - tri[0])
        assert math.sqrt(brute) - math.sqrt(d2) <= h / 300.0 + 1e-9


def _box_tris(lo, hi):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c = np.array([[lo[0] if (i & 1) == 0 else hi[0], lo[1] if (i & 2) == 0 else hi[1],
                   lo[2] if (i & 4) == 0 else hi[2]] for i in range(8)])
    faces = [(0, 4, 6, 2), (1, 3, 7, 5), (0, 1, 5, 4), (2, 6, 7, 3), (0, 2, 3, 1), (4, 5, 7, 6)]
    tris = []
    for a, b, cc, d in faces:
        tris.append(np.concatenate([c[a], c[b], c[cc]]))
        tris.append(np.concatenate([c[a], c[cc], c[d]]))
    return np.array(tris, dtype=np.float32)


def test_box_mesh_is_outward(oracle):
    tris = _box_tris((1.0, 2.0, 3.0), (4.0, 6.0, 8.0))
    assert signed_volume(tris) == pytest.approx(3 * 4 * 5)


def test_make_sdf_box_matches_analytic(oracle):
    cfg = config(oracle, SIZE_X=16, SIZE_Y=16, SIZE_Z=16)
    lo, hi = np.array([4.3, 3.7, 5.2]), np.array([11.6, 9.4, 10.1])
    sdf = oracle.make_sdf(cfg, _box_tris(lo, hi), nthreads=4)
    X = node_coords(cfg)
    c, h = (lo + hi) / 2, (hi - lo) / 2
    q = np.abs(X - c[:, None]) - h[:, None]
    ana = np.linalg.norm(np.maximum(q, 0), axis=0) + np.minimum(q.max(axis=0), 0)
    np.testing.assert_allclose(sdf, ana, atol=2e-5)


def test_make_sdf_sphere_mesh_within_sagitta(oracle, mpm):
    cfg = config(oracle, SIZE_X=24, SIZE_Y=24, SIZE_Z=24)
    ctr, r = np.array([11.7, 12.2, 11.9]), 7.0
    tris = mpm.mesh_uv_sphere(ctr, r, 32, 33)
    sdf = oracle.make_sdf(cfg, tris, nthreads=4)
    X = node_coords(cfg)
    ana = np.linalg.norm(X - ctr[:, None], axis=0) - r
    # a triangle of a (2*pi/32) x (pi/33) quad lies within angle rho of its
    # vertices, so it sits at most r*(1 - cos(rho)) inside the sphere
    rho = 0.5 * math.hypot(2 * math.pi / 32, math.pi / 33)
    sagitta = r * (1 - math.cos(rho))
    assert np.abs(sdf - ana).max() <= sagitta + 1e-4
    far = np.abs(ana) > sagitta + 1e-3
    np.testing.assert_array_equal(np.sign(sdf[far]), np.sign(ana[far]))


def test_make_sdf_no_triangles(oracle):
    cfg = config(oracle, SIZE_X=4, SIZE_Y=4, SIZE_Z=4)
    sdf = oracle.make_sdf(cfg, np.zeros((0, 9), np.float32))
    np.testing.assert_allclose(sdf, 1e15, rtol=1e-6)
