"""GPU tests of the batched engine, HIP-graph replay, snapshot/restore and the
autotuner (build additions with no reference counterpart).

A batch must reproduce, scene by scene, what the single-scene engine and the
oracle compute for that scene alone: the scenes share launches, never data.
Tolerances are those of the single-scene parity tests (float-atomic order).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POS_TOL = 1e-4
VEL_TOL = 1e-4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def max_err(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) if a.size else 0.0


@pytest.fixture(scope="module")
def small_batch(gpu):
    # Six scenes of different grid sizes and particle counts (some scenes
    # share nothing but launches with much larger ones).
    return gpu.scenes.batch_scenes(count=6, seed=11, n_min=2000, n_max=60000, edges=(32, 48, 64))


def batch_engine(mpm, b, **opts):
    e = mpm.Batch(b["cfgs"], **opts)
    e.init(b["counts"], b["particles"])
    for s, sdf in enumerate(b["sdfs"]):
        e.set_scene_sdf(s, sdf)
    return e


def scene_slice(b, s):
    lo = int(b["counts"][:s].sum())
    return lo, lo + int(b["counts"][s])


def test_batch_matches_single_scene_engines(gpu, small_batch):
    b = small_batch
    say(f"batch: {len(b['cfgs'])} scenes, n={int(b['counts'].sum())}, one substep and 12 substeps")
    eb = batch_engine(gpu, b, RESORT_EVERY=5)
    eb.step()
    got1 = eb.getParticles()
    eb.steps(11)
    got12 = eb.getParticles()
    for s, cfg in enumerate(b["cfgs"]):
        lo, hi = scene_slice(b, s)
        e = gpu.Engine(cfg, RESORT_EVERY=5)
        e.init(hi - lo, np.ascontiguousarray(b["particles"][:, lo:hi]))
        e.set_sdf(b["sdfs"][s])
        e.step()
        one = e.getParticles()
        e.steps(11)
        twelve = e.getParticles()
        e.close()
        assert max_err(got1[0:3, lo:hi], one[0:3]) < POS_TOL, s
        assert max_err(got1[3:6, lo:hi], one[3:6]) < VEL_TOL, s
        assert max_err(got12[0:3, lo:hi], twelve[0:3]) < 12 * POS_TOL, s
        assert max_err(got12[3:6, lo:hi], twelve[3:6]) < 12 * VEL_TOL, s
        np.testing.assert_array_equal(got12[24:26, lo:hi], twelve[24:26])
    eb.close()


def test_odd_size_batch_matches_single_scene_engines(gpu):
    """Scenes whose edges (21, 37, 45) give block counts that are not
    multiples of 4: the block numbering, the chunk table and the grid
    update's row segments cross scene boundaries in the middle of a segment
    and with partial blocks."""
    b = gpu.scenes.batch_scenes(count=5, seed=23, n_min=2000, n_max=40000, edges=(21, 37, 45))
    say(f"odd batch: edges {[c.SIZE_X for c in b['cfgs']]}, n={int(b['counts'].sum())}")
    eb = batch_engine(gpu, b, RESORT_EVERY=3)
    eb.steps(8)
    got = eb.getParticles()
    for s, cfg in enumerate(b["cfgs"]):
        lo, hi = scene_slice(b, s)
        e = gpu.Engine(cfg, RESORT_EVERY=3)
        e.init(hi - lo, np.ascontiguousarray(b["particles"][:, lo:hi]))
        e.set_sdf(b["sdfs"][s])
        e.steps(8)
        ref = e.getParticles()
        e.close()
        assert max_err(got[0:3, lo:hi], ref[0:3]) < 8 * POS_TOL, s
        assert max_err(got[3:6, lo:hi], ref[3:6]) < 8 * VEL_TOL, s
    eb.close()


def test_batch_matches_oracle(gpu, oracle, small_batch):
    b = small_batch
    say("batch vs oracle: one substep per scene")
    eb = batch_engine(gpu, b)
    eb.step()
    got = eb.getParticles()
    for s, cfg in enumerate(b["cfgs"]):
        lo, hi = scene_slice(b, s)
        st = oracle.init_state(cfg, np.ascontiguousarray(b["particles"][:, lo:hi]), np.float32)
        ref = oracle.steps(cfg, st, b["sdfs"][s], 1)
        assert max_err(got[0:3, lo:hi], ref[0:3]) < POS_TOL, s
        assert max_err(got[3:6, lo:hi], ref[3:6]) < VEL_TOL, s
    eb.close()


def test_batch_scene_grid_and_info(gpu, small_batch):
    b = small_batch
    eb = batch_engine(gpu, b)
    eb.steps(2)
    whole = eb.getGrid()
    total = 0
    for s, cfg in enumerate(b["cfgs"]):
        info = eb.scene_info(s)
        assert info["count"] == int(b["counts"][s])
        assert info["nodes"] == cfg.nodes
        g = eb.scene_grid(s)
        np.testing.assert_array_equal(g, whole[:, info["node_offset"]:info["node_offset"] + info["nodes"]])
        # mass of the scene's grid equals its particle count (m = 1) when no
        # particle sits within two cells of the domain boundary
        total += float(g[0].sum())
    assert abs(total - float(b["counts"].sum())) < 1e-3 * float(b["counts"].sum())
    eb.close()


def test_graph_replay_matches_eager(gpu, small_batch):
    b = small_batch
    say("graph: 3 x steps(10) captured once and replayed, vs eager")
    a = batch_engine(gpu, b, RESORT_EVERY=5)
    g = batch_engine(gpu, b, RESORT_EVERY=5, GRAPH=1)
    for call in range(3):
        a.steps(10)
        g.steps(10)
        # eager work between replays
        pa, pg = a.getParticles(), g.getParticles()
        assert max_err(pa[0:3], pg[0:3]) < 10 * (call + 1) * POS_TOL, call
        assert max_err(pa[3:6], pg[3:6]) < 10 * (call + 1) * VEL_TOL, call
    assert g.get_option("GRAPH_REPLAYS") == 3
    a.close(); g.close()


def test_graph_single_scene_sphere(gpu):
    sc = gpu.scenes.sphere64(physical=True)
    say("graph: sphere64, steps(7) x 4 with a 3-substep re-sort period")
    a = gpu.Engine(sc["cfg"], RESORT_EVERY=3)
    g = gpu.Engine(sc["cfg"], RESORT_EVERY=3, GRAPH=1)
    for e in (a, g):
        e.init(sc["particles"].shape[1], sc["particles"])
        e.set_sdf(sc["sdf"])
    for call in range(4):
        a.steps(7)
        g.steps(7)
        pa, pg = a.getParticles(), g.getParticles()
        assert max_err(pa[0:3], pg[0:3]) < 7 * (call + 1) * POS_TOL
    a.close(); g.close()


def test_snapshot_restore_rewinds(gpu, small_batch):
    b = small_batch
    e = batch_engine(gpu, b, RESORT_EVERY=4)
    e.steps(3)
    e.snapshot()
    before = e.getParticles()
    e.steps(6)
    first = e.getParticles()
    e.restore()
    np.testing.assert_array_equal(e.getParticles(), before)
    e.set_option("CHUNK", 256)  # another layout: same physics
    e.steps(6)
    second = e.getParticles()
    assert max_err(first[0:3], second[0:3]) < 6 * POS_TOL
    e.close()


def test_autotune_keeps_state_and_caches(gpu, small_batch, tmp_path):
    from importlib import import_module

    at = import_module("water-sim-cuda_amd.autotune")
    b = small_batch
    e = batch_engine(gpu, b, RESORT_EVERY=5)
    e.steps(2)
    before = e.getParticles()
    cache = tmp_path / "tune.json"
    res = at.autotune(e, trial_substeps=3, warmup=1, chunks=(128, 512), launch_pcts=(100, 50), cache=cache)
    say(f"autotune: {res['CHUNK']} / {res['LAUNCH_PCT']} ({len(res['trials'])} trials)")
    assert not res["cached"] and len(res["trials"]) == 4
    assert e.get_option("CHUNK") == res["CHUNK"]
    np.testing.assert_array_equal(e.getParticles(), before)
    res2 = at.autotune(e, cache=cache)
    assert res2["cached"] and res2["CHUNK"] == res["CHUNK"]
    e.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_online_tuner_runs_the_frames(gpu, small_batch, graph):
    """Trial frames are frames of the run: after twelve 2-substep frames with
    a trial in every fourth frame, the state is that of 24 plain substeps
    (within float-atomic order) whatever CHUNK the tuner kept.  With graph
    replay, frames replay their graph, a CHUNK change makes the next
    (untimed) frame capture a new one, and no timed frame captures."""
    from importlib import import_module

    at = import_module("water-sim-cuda_amd.autotune")
    b = small_batch
    ref = batch_engine(gpu, b, RESORT_EVERY=2, CHUNK=256)
    ref.steps(24)
    expect = ref.getParticles()
    ref.close()
    e = batch_engine(gpu, b, RESORT_EVERY=2, CHUNK=256, GRAPH=graph)
    tuner = at.OnlineTuner(e, frame_substeps=2, period=4, chunks=(128, 256, 512))
    for _ in range(12):
        tuner.frame()
    say(f"online tuner (graph {graph}): {tuner.trials} trials, {tuner.switches} switches, "
        f"CHUNK {e.get_option('CHUNK')}, history {tuner.history}, replays {e.get_option('GRAPH_REPLAYS')}, "
        f"captures {e.get_option('GRAPH_CAPTURES')}")
    assert tuner.frames == 12 and len(tuner.history) == tuner.trials
    assert e.get_option("CHUNK") in (128, 256, 512)
    if graph:
        # every frame ran as a graph; later frames of one setting replay
        assert e.get_option("GRAPH_REPLAYS") == 12
        assert e.get_option("GRAPH_CAPTURES") < 12
        assert tuner.trials >= 1
    else:
        assert tuner.trials == 3 and not any(h[5] for h in tuner.history)
    got = e.getParticles()
    assert np.isfinite(got).all()
    assert max_err(got[0:3], expect[0:3]) < 24 * POS_TOL
    assert max_err(got[3:6], expect[3:6]) < 24 * VEL_TOL
    e.close()


def test_resort_every_change_drops_the_graph(gpu, small_batch):
    """Changing RESORT_EVERY under graph replay takes effect: the next call
    captures a new graph, and the state follows an eager engine run with the
    same schedule (within float-atomic order)."""
    b = small_batch
    eager = batch_engine(gpu, b, RESORT_EVERY=4, CHUNK=256)
    e = batch_engine(gpu, b, RESORT_EVERY=4, CHUNK=256, GRAPH=1)
    for _ in range(6):  # until the start states of consecutive calls repeat
        c = e.get_option("GRAPH_CAPTURES")
        e.steps(4)
        eager.steps(4)
        if e.get_option("GRAPH_CAPTURES") == c:
            break
    c0 = e.get_option("GRAPH_CAPTURES")
    for _ in range(2):
        e.steps(4)
        eager.steps(4)
    assert e.get_option("GRAPH_CAPTURES") == c0  # same schedule: replays
    for eng in (eager, e):
        eng.set_option("RESORT_EVERY", 1)
    e.steps(4)
    assert e.get_option("GRAPH_CAPTURES") == c0 + 1  # new schedule: new graph
    eager.steps(4)
    got, expect = e.getParticles(), eager.getParticles()
    assert max_err(got[0:3], expect[0:3]) < 40 * POS_TOL
    assert max_err(got[3:6], expect[3:6]) < 40 * VEL_TOL
    e.close()
    eager.close()
