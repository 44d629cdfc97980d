"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Tolerances.  Morton keys and the sort permutation are integer results and
must match bit for bit.  Float results differ from the oracle by float-atomic
summation order and fused multiply-adds only: one substep must agree within
1e-4 cell widths on positions and 1e-4 cells/s on velocities (the bar the
project sets for FP32), stage outputs within the tolerances given per test.
"""
import sys

import numpy as np
import pytest

from helpers import keys_numpy, node_coords, stable_perm

pytestmark = pytest.mark.gpu

POS_TOL = 1e-4  # cell widths, one substep
VEL_TOL = 1e-4  # cells per second, one substep


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def engine_for(mpm, scene, **opts):
    e = mpm.Engine(scene["cfg"], **opts)
    e.init(scene["particles"].shape[1], scene["particles"])
    if scene.get("sdf") is not None:
        e.set_sdf(scene["sdf"])
    else:
        e.initSDF(scene["triangles"])
    return e


def oracle_state(oracle, scene):
    return oracle.init_state(scene["cfg"], scene["particles"], np.float32)


def max_err(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) if a.size else 0.0


@pytest.fixture(scope="module")
def sphere(gpu):
    return gpu.scenes.sphere64(physical=False)


@pytest.fixture(scope="module")
def sphere_phys(gpu):
    return gpu.scenes.sphere64(physical=True)


# ---------------- sort (bit-exact) ----------------

def test_sort_bitexact_sphere64(gpu, oracle, sphere):
    say("sort: sphere64")
    e = engine_for(gpu, sphere)
    keys, perm = e.sortParticles()
    pos = sphere["particles"][0:3]
    okeys = oracle.morton_keys(pos)
    operm = oracle.sort_perm(okeys)
    np.testing.assert_array_equal(perm, operm)
    np.testing.assert_array_equal(keys, okeys[operm])
    np.testing.assert_array_equal(okeys, keys_numpy(pos))
    e.close()


def test_sort_bitexact_out_of_domain(gpu, oracle):
    """Keys of positions below zero and beyond the grid, and many ties."""
    say("sort: out-of-domain positions")
    cfg = gpu.make_config(SIZE_X=64, SIZE_Y=64, SIZE_Z=64)
    rng = np.random.default_rng(10)
    n = 20000
    p = np.zeros((26, n), dtype=np.float32)
    p[0:3] = rng.uniform(-10.0, 700.0, size=(3, n))
    p[0:3, :5000] = np.floor(rng.uniform(0, 4, size=(3, 5000))) + 0.25  # ties
    e = gpu.Engine(cfg)
    e.init(n, p)
    keys, perm = e.sortParticles()
    okeys = oracle.morton_keys(p[0:3])
    operm = oracle.sort_perm(okeys)
    np.testing.assert_array_equal(perm, operm)
    np.testing.assert_array_equal(keys, okeys[operm])
    e.close()


# ---------------- stages ----------------

def test_p2g_stage(gpu, oracle, sphere):
    say("stage: p2g sphere64")
    e = engine_for(gpu, sphere)
    g = e.debug_p2g()
    st = oracle_state(oracle, sphere)
    ref = oracle.p2g(sphere["cfg"], st)
    assert float(g[0].astype(np.float64).sum()) == pytest.approx(st.shape[1], rel=1e-5)
    np.testing.assert_allclose(g, ref, rtol=1e-4, atol=2e-5)
    # state untouched by the stage call
    np.testing.assert_array_equal(e.getParticles()[0:24], st[0:24])
    e.close()


def test_grid_update_stage(gpu, oracle):
    say("stage: grid update random grid")
    cfg = gpu.make_config(SIZE_X=40, SIZE_Y=36, SIZE_Z=28, PADDING=3, GRAVITY=9.8, DT=1e-3)
    X = node_coords(cfg)
    rng = np.random.default_rng(11)
    G = cfg.nodes
    grid = np.zeros((4, G), dtype=np.float32)
    grid[0] = rng.uniform(0.0, 3.0, size=G) * (rng.uniform(size=G) < 0.7)
    grid[0, ::97] = 5e-10
    grid[1:4] = rng.normal(size=(3, G)) * 4.0
    sdf = np.minimum(X[1] - 9.3, 14.2 - np.abs(X[0] - 20.0)).astype(np.float32)
    e = gpu.Engine(cfg)
    e.init(0, np.zeros(0, np.float32))
    e.set_sdf(sdf)
    out = e.debug_grid_update(grid)
    ref = oracle.grid_update(cfg, grid, sdf)
    np.testing.assert_allclose(out, ref, rtol=2e-6, atol=2e-6)
    e.close()


def test_g2p_stage(gpu, oracle, sphere_phys):
    say("stage: g2p sphere64")
    sc = sphere_phys
    cfg = sc["cfg"]
    st = oracle_state(oracle, sc)
    st[3:6] = np.random.default_rng(12).normal(size=(3, st.shape[1])).astype(np.float32)
    grid = oracle.grid_update(cfg, oracle.p2g(cfg, st), sc["sdf"])
    e = gpu.Engine(cfg)
    e.init(st.shape[1], st)
    e.set_sdf(sc["sdf"])
    e.debug_g2p(grid)
    got = e.getParticles()
    ref = oracle.g2p(cfg, st, grid, sc["sdf"])
    assert max_err(got[0:3], ref[0:3]) < 1e-5
    assert max_err(got[3:6], ref[3:6]) < 1e-5
    assert max_err(got[6:15], ref[6:15]) < 1e-4
    assert max_err(got[15:24], ref[15:24]) < 1e-6
    e.close()


# ---------------- whole substeps ----------------

@pytest.mark.parametrize("physical", [False, True])
def test_one_substep_sphere64(gpu, oracle, physical):
    sc = gpu.scenes.sphere64(physical=physical)
    say(f"substep: sphere64 physical={physical} n={sc['particles'].shape[1]}")
    e = engine_for(gpu, sc)
    e.step()
    got = e.getParticles()
    st = oracle_state(oracle, sc)
    ref = oracle.steps(sc["cfg"], st, sc["sdf"], 1)
    ref64 = oracle.steps(sc["cfg"], oracle.init_state(sc["cfg"], sc["particles"], np.float64),
                         sc["sdf"].astype(np.float64), 1)
    pos_err, vel_err = max_err(got[0:3], ref[0:3]), max_err(got[3:6], ref[3:6])
    say(f"  gpu-vs-fp32 pos {pos_err:.3e} vel {vel_err:.3e}; fp32-vs-fp64 pos "
        f"{max_err(ref[0:3], ref64[0:3]):.3e} vel {max_err(ref[3:6], ref64[3:6]):.3e}")
    assert pos_err < POS_TOL
    assert vel_err < VEL_TOL
    assert max_err(got[6:15], ref[6:15]) < 1e-3
    assert max_err(got[15:24], ref[15:24]) < 1e-5
    np.testing.assert_array_equal(got[24:26], ref[24:26])
    e.close()


def test_twenty_substeps_sphere64(gpu, oracle, sphere_phys):
    sc = sphere_phys
    say("substeps: sphere64 x20")
    e = engine_for(gpu, sc)
    e.steps(20)
    got = e.getParticles()
    ref = oracle.steps(sc["cfg"], oracle_state(oracle, sc), sc["sdf"], 20, nthreads=8)
    # 20 substeps accumulate the per-substep rounding differences
    assert max_err(got[0:3], ref[0:3]) < 20 * POS_TOL
    assert max_err(got[3:6], ref[3:6]) < 20 * VEL_TOL
    e.close()


def test_call_boundaries_keep_the_substeps(gpu, oracle, sphere_phys):
    """A steps() call ends with a fused transfer that also runs the next
    substep's P2G into the grid, and the next call starts from it: calls of
    any size, and a CHUNK change (a new layout) between calls, give the
    oracle's 16 substeps."""
    sc = sphere_phys
    say("substeps: call boundaries (8+8, 16 x step(), 3+CHUNK+5+8)")
    ref = oracle.steps(sc["cfg"], oracle_state(oracle, sc), sc["sdf"], 16, nthreads=8)
    a = engine_for(gpu, sc)
    a.steps(8)
    a.steps(8)
    b = engine_for(gpu, sc)
    for _ in range(16):
        b.step()
    c = engine_for(gpu, sc, CHUNK=2048)
    c.steps(3)
    c.set_option("CHUNK", 256)
    c.steps(5)
    c.steps(8)
    for name, e in (("8+8", a), ("16 x step()", b), ("3+CHUNK+5+8", c)):
        got = e.getParticles()
        err = max_err(got[0:3], ref[0:3])
        say(f"  {name}: position error {err:.2e}")
        assert err < 16 * POS_TOL
        assert max_err(got[3:6], ref[3:6]) < 16 * VEL_TOL
        e.close()


def test_pending_p2g_dropped_when_state_changes(gpu, oracle, sphere_phys):
    """The next substep's P2G that a call leaves in the grid is dropped when
    the particles or DT change outside steps(): after restore() and after
    load_state() the run continues from the restored particles, and after a
    DT change the next substep uses the new DT."""
    sc = sphere_phys
    say("substeps: pending P2G after restore, load_state and a DT change")
    s0 = oracle_state(oracle, sc)
    ref4 = oracle.steps(sc["cfg"], s0, sc["sdf"], 4, nthreads=8)
    e = engine_for(gpu, sc)
    e.snapshot()
    e.steps(6)
    e.restore()
    e.steps(4)
    assert max_err(e.getParticles()[0:3], ref4[0:3]) < 4 * POS_TOL
    e.steps(3)
    e.load_state(s0)
    e.steps(4)
    assert max_err(e.getParticles()[0:3], ref4[0:3]) < 4 * POS_TOL
    e.close()
    cfg2 = gpu.SimConfig.from_buffer_copy(sc["cfg"])
    cfg2.DT = sc["cfg"].DT * 2.0
    ref = oracle.steps(cfg2, oracle.steps(sc["cfg"], s0, sc["sdf"], 3, nthreads=8), sc["sdf"], 3, nthreads=8)
    f = engine_for(gpu, sc)
    f.steps(3)
    f.set_dt_scale(2.0)
    f.steps(3)
    err = max_err(f.getParticles()[0:3], ref[0:3])
    say(f"  DT change: position error {err:.2e}")
    assert err < 6 * POS_TOL
    f.close()


def test_fused_and_single_steps_agree(gpu, sphere_phys):
    """steps(k) runs the fused G2P+P2G kernel; k calls of step() do not."""
    say("substeps: fused vs single")
    a = engine_for(gpu, sphere_phys)
    b = engine_for(gpu, sphere_phys)
    a.steps(8)
    for _ in range(8):
        b.step()
    pa, pb = a.getParticles(), b.getParticles()
    assert max_err(pa[0:3], pb[0:3]) < 8 * POS_TOL
    assert max_err(pa[3:6], pb[3:6]) < 8 * VEL_TOL
    a.close(); b.close()


def test_resort_period_does_not_change_results(gpu, sphere_phys):
    say("substeps: resort every 1 vs 7")
    a = engine_for(gpu, sphere_phys, RESORT_EVERY=1)
    b = engine_for(gpu, sphere_phys, RESORT_EVERY=7)
    a.steps(10); b.steps(10)
    pa, pb = a.getParticles(), b.getParticles()
    assert max_err(pa[0:3], pb[0:3]) < 10 * POS_TOL
    a.close(); b.close()


@pytest.mark.parametrize("chunk", [64, 256, 1024])
def test_chunk_sizes(gpu, oracle, sphere_phys, chunk):
    say(f"substep: chunk {chunk}")
    e = engine_for(gpu, sphere_phys, CHUNK=chunk)
    e.steps(2)
    got = e.getParticles()
    ref = oracle.steps(sphere_phys["cfg"], oracle_state(oracle, sphere_phys), sphere_phys["sdf"], 2)
    assert max_err(got[0:3], ref[0:3]) < 2 * POS_TOL
    e.close()


def test_out_of_tile_particles(gpu, oracle):
    """Fast particles leave their workgroup's LDS tile between re-sorts and
    take the global-atomic path; results still match."""
    say("substeps: tile overflow path")
    sc = gpu.scenes.small_box(n_cells=(8, 8, 8), grid=48, physical=True, dt=2e-3)
    sc["particles"][3] = 300.0  # 0.6 cells per substep along x
    sc["particles"][5] = -150.0
    e = engine_for(gpu, sc, RESORT_EVERY=50)
    e.steps(6)
    assert e.counters()["tile_overflow_particles"] > 0
    got = e.getParticles()
    ref = oracle.steps(sc["cfg"], oracle_state(oracle, sc), sc["sdf"], 6)
    assert max_err(got[0:3], ref[0:3]) < 6 * POS_TOL
    e.close()


def test_particles_outside_grid(gpu, oracle):
    """Particles below 0 and beyond SIZE (no in-bounds nodes) and at the edge."""
    say("substep: particles outside the grid")
    cfg = gpu.make_config(SIZE_X=32, SIZE_Y=32, SIZE_Z=32, GRAVITY=9.8, GAMMA=-7.0)
    rng = np.random.default_rng(13)
    n = 3000
    p = np.zeros((26, n), dtype=np.float32)
    p[0:3] = rng.uniform(-6.0, 38.0, size=(3, n))
    p[3:6] = rng.normal(size=(3, n))
    p[24:26] = 1.0
    sdf = np.full(cfg.nodes, 1e15, dtype=np.float32)
    e = gpu.Engine(cfg)
    e.init(n, p)
    e.set_sdf(sdf)
    e.step()
    got = e.getParticles()
    ref = oracle.steps(cfg, oracle.init_state(cfg, p, np.float32), sdf, 1)
    assert max_err(got[0:3], ref[0:3]) < POS_TOL
    assert max_err(got[3:6], ref[3:6]) < VEL_TOL
    e.close()


def test_zero_and_one_particle(gpu):
    say("edge: n=0 and n=1")
    cfg = gpu.make_config(SIZE_X=32, SIZE_Y=32, SIZE_Z=32, GRAVITY=9.8, DT=1e-3)
    e = gpu.Engine(cfg)
    e.init(0, np.zeros(0, np.float32))
    e.steps(3)
    assert e.getNumber() == 0
    p = np.zeros((26, 1), dtype=np.float32)
    p[0:3, 0] = (15.3, 16.7, 14.2)
    p[3:6, 0] = (0.3, 1.0, -0.2)
    e.init(1, p)
    e.steps(3)
    got = e.getParticles()[:, 0]
    gdt = np.float32(9.8) * np.float32(1e-3)
    v = np.array([0.3, 1.0 - 3 * gdt, -0.2])
    assert np.abs(got[3:6] - v).max() < 1e-5
    e.close()


def test_positions_export(gpu, sphere_phys):
    e = engine_for(gpu, sphere_phys)
    e.steps(2)
    xyz = e.getPositions()
    np.testing.assert_array_equal(xyz.T, e.getParticles()[0:3])
    # the same read-back into page-locked host memory
    pin = gpu.pinned_array(xyz.shape)
    assert e.getPositions(out=pin) is pin
    np.testing.assert_array_equal(pin, xyz)
    with pytest.raises(ValueError):
        e.getPositions(out=np.empty((3, xyz.shape[0]), np.float32))
    e.close()


@pytest.mark.parametrize("dims", [(37, 29, 41), (45, 22, 19)])
def test_odd_grid_substeps(gpu, oracle, dims):
    """Grid edges that are not multiples of 4 or 16: partial 4^3 blocks at the
    far faces and partial row segments of four blocks in the grid update.
    The water fills the grid up to the far walls with a swirling velocity
    field (particles cross cells and blocks and reach the walls within the
    run), above a box collider (mesh SDF), re-sorted every substep and every
    third substep."""
    sx, sy, sz = dims
    say(f"substeps: odd grid {dims}")
    cfg = gpu.make_config(SIZE_X=sx, SIZE_Y=sy, SIZE_Z=sz, PADDING=3, DT=4e-3, GRAVITY=9.8, GAMMA=-7.0,
                          COMPRESSION=10.0)
    parts = gpu.scene_fill("box", (3.5, 5.0, 3.5), (sx - 3.5, sy - 4.0, sz - 3.5), 11)
    x, y, z = parts[0], parts[1], parts[2]
    parts[3] = (30.0 * np.sin(y / 4.0)).astype(np.float32)
    parts[4] = (20.0 * np.cos(x / 5.0) * np.sin(z / 6.0)).astype(np.float32)
    parts[5] = (25.0 * np.sin(x / 3.0 + y / 7.0)).astype(np.float32)
    tris = gpu.mesh_box((5.2, 1.0, 6.1), (sx - 6.3, 4.6, sz - 5.7))
    sdf = None
    for every in (1, 3):
        e = gpu.Engine(cfg, RESORT_EVERY=every)
        e.init(parts.shape[1], parts)
        e.initSDF(tris)
        if sdf is None:
            sdf = e.get_sdf()
            ref = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, 6, nthreads=8)
            moved = float(np.abs(ref[0:3] - parts[0:3]).max())
            say(f"  largest displacement {moved:.2f} cells")
        e.steps(6)
        got = e.getParticles()
        ep, ev = max_err(got[0:3], ref[0:3]), max_err(got[3:6], ref[3:6])
        say(f"  {parts.shape[1]} particles, RESORT_EVERY {every}: position error {ep:.2e}, velocity error {ev:.2e}")
        assert ep < 6 * POS_TOL and ev < 60 * VEL_TOL, every
        e.close()


# ---------------- mesh SDF ----------------

@pytest.mark.parametrize("mesh", ["sphere", "box", "channel"])
def test_make_sdf_parity(gpu, oracle, mesh):
    say(f"makeSDF: {mesh}")
    cfg = gpu.make_config(SIZE_X=28, SIZE_Y=24, SIZE_Z=26)
    if mesh == "sphere":
        tris = gpu.mesh_uv_sphere((13.3, 11.8, 12.6), 8.0, 32, 33)
    elif mesh == "box":
        tris = gpu.mesh_box((4.3, 3.7, 5.2), (21.6, 9.4, 17.1))
    else:
        tris = gpu.mesh_u_channel((2.3, 3.3, 2.3), (25.7, 15.3, 23.7), 3.0, 4.0)
    e = gpu.Engine(cfg)
    e.initSDF(tris)
    got = e.get_sdf()
    ref = oracle.make_sdf(cfg, tris, nthreads=8)
    far = np.abs(ref) > 1e-3
    np.testing.assert_array_equal(np.sign(got[far]), np.sign(ref[far]))
    assert np.abs(np.abs(got) - np.abs(ref)).max() < 1e-4
    e.close()


def test_no_triangles_gives_far_field(gpu):
    cfg = gpu.make_config(SIZE_X=8, SIZE_Y=8, SIZE_Z=8)
    e = gpu.Engine(cfg)
    e.initSDF(np.zeros((0, 9), np.float32))
    np.testing.assert_allclose(e.get_sdf(), 1e15, rtol=1e-6)
    e.close()
