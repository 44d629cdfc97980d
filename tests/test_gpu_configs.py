"""The benchmark workloads at their full sizes (GPU), against the oracle.

Each workload runs through the C ABI at the size it is benchmarked at:

* sphere64   -- 64^3, water sphere r = 12 (57.9 k particles), the literal
                parameters GRAVITY = -9.8, GAMMA = 7; its whole run of 100
                frames x 20 substeps on the GPU and on the oracle.
* dambreak128 -- 128^3 dam break (1,769,472 particles), UV-sphere collider,
                re-sort every 10 substeps; its whole run of 200 frames x 20.
* channel256 -- 256^3 (11,534,336 particles), U-channel collider, one frame
                of 40 substeps.
* batch64    -- 64 scenes (edges 32-96, 5 k-400 k particles), graph replay
                and autotuned CHUNK, 100 frames x 20 substeps.

For each, one substep is compared particle by particle with the FP32 oracle
(positions within 1e-4 cells, velocities within 1e-4 cells/s: float-atomic
order is the only difference) and with the FP64 oracle, whose gap to the FP32
oracle is the rounding floor of the formulas themselves; the three gaps go
to $MPM_PARITY_LOG as one JSON line per workload.  Longer runs are checked
by what a correct step keeps: every particle finite, the grid's mass equal
to the particle count (m = 1 each), no particle deeper than half a cell in
the collider, and for whole runs agreement of height and density statistics
with the oracle (sphere64) or between re-sort periods (dambreak128; the
re-sort changes only the order of float sums, engine.cu:57).
"""
import sys
import time

import numpy as np
import pytest

from helpers import host_threads, parity_log

pytestmark = pytest.mark.gpu

POS_TOL = 1e-4  # cell widths, one substep
VEL_TOL = 1e-4  # cells per second, one substep


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def max_err(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) if a.size else 0.0


def engine_for(mpm, scene, **opts):
    e = mpm.Engine(scene["cfg"], **opts)
    e.init(scene["particles"].shape[1], scene["particles"])
    if scene.get("sdf") is not None:
        e.set_sdf(scene["sdf"])
    else:
        e.initSDF(scene["triangles"])
    return e


def one_substep_gaps(oracle, cfg, parts, sdf, got) -> dict:
    """Max |gap| of positions and velocities after one substep: GPU vs FP32
    oracle, GPU vs FP64 oracle, FP32 vs FP64 oracle."""
    th = host_threads()
    ref32 = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, 1, nthreads=th)
    ref64 = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float64), sdf.astype(np.float64), 1, nthreads=th)
    return {
        "gpu_vs_fp32": {"pos": max_err(got[0:3], ref32[0:3]), "vel": max_err(got[3:6], ref32[3:6])},
        "gpu_vs_fp64": {"pos": max_err(got[0:3], ref64[0:3]), "vel": max_err(got[3:6], ref64[3:6])},
        "fp32_vs_fp64": {"pos": max_err(ref32[0:3], ref64[0:3]), "vel": max_err(ref32[3:6], ref64[3:6])},
    }


def check_one_substep(name, gaps, n):
    say(f"{name}: one substep, n={n}: {gaps}")
    parity_log({"workload": name, "particles": n, **gaps})
    assert gaps["gpu_vs_fp32"]["pos"] < POS_TOL and gaps["gpu_vs_fp32"]["vel"] < VEL_TOL
    assert gaps["gpu_vs_fp64"]["pos"] < POS_TOL and gaps["gpu_vs_fp64"]["vel"] < VEL_TOL


def stats(state):
    pos = state[0:3].astype(np.float64)
    y = pos[1]
    cells = np.floor(pos).astype(np.int64)
    key = (cells[2] * 4096 + cells[1]) * 4096 + cells[0]
    _, counts = np.unique(key, return_counts=True)
    return {"mean_y": float(y.mean()), "std_y": float(y.std()), "p05_y": float(np.percentile(y, 5)),
            "p95_y": float(np.percentile(y, 95)), "occupied": int(len(counts)),
            "mean_per_cell": float(counts.mean()),
            "mean_speed": float(np.sqrt((state[3:6].astype(np.float64) ** 2).sum(0)).mean())}


def stat_gaps(a, b):
    return {"mean_y": abs(a["mean_y"] - b["mean_y"]), "std_y": abs(a["std_y"] - b["std_y"]),
            "p05_y": abs(a["p05_y"] - b["p05_y"]), "p95_y": abs(a["p95_y"] - b["p95_y"]),
            "occupied": abs(a["occupied"] - b["occupied"]) / b["occupied"],
            "mean_per_cell": abs(a["mean_per_cell"] - b["mean_per_cell"]) / b["mean_per_cell"],
            "mean_speed": abs(a["mean_speed"] - b["mean_speed"]) / max(b["mean_speed"], 1e-9)}


def check_invariants(name, eng, n, grid_mass):
    st = eng.getParticles()
    chk = eng.check(0.5)
    say(f"{name}: grid mass {grid_mass:.3f} (n={n}), check {chk}")
    assert np.isfinite(st[0:24]).all()
    assert abs(grid_mass - n) < 1e-3 * n
    assert chk["nonfinite"] == 0 and chk["penetrating"] == 0
    return st


# ---------------- sphere64: the whole run against the oracle ----------------

def test_sphere64_one_substep_and_whole_run(gpu, oracle):
    sc = gpu.scenes.sphere64(physical=False)  # GRAVITY = -9.8, GAMMA = 7 as stated for this workload
    cfg, parts, sdf = sc["cfg"], sc["particles"], sc["sdf"]
    n = parts.shape[1]
    e = engine_for(gpu, sc)
    e.step()
    check_one_substep("sphere64", one_substep_gaps(oracle, cfg, parts, sdf, e.getParticles()), n)
    e.close()

    frames, k = sc["frames"], int(cfg.SUBSTEPS)
    runs = {}
    for resort in (1, 7):
        e = engine_for(gpu, sc, RESORT_EVERY=resort)
        for _ in range(frames):
            e.steps(k)
        runs[resort] = check_invariants(f"sphere64 resort {resort}", e, n, float(e.getGrid()[0].astype(np.float64).sum()))
        e.close()
    t0 = time.time()
    ref = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, frames * k, nthreads=host_threads())
    say(f"sphere64: oracle {frames * k} substeps in {time.time() - t0:.1f}s")
    sg, so, s7 = stats(runs[1]), stats(ref), stats(runs[7])
    d, d_self = stat_gaps(sg, so), stat_gaps(s7, sg)
    say(f"sphere64 after {frames} frames:\n  gpu {sg}\n  oracle {so}\n  gpu-vs-oracle {d}\n  gpu-vs-gpu {d_self}")
    parity_log({"workload": "sphere64", "whole_run_substeps": frames * k, "gpu": sg, "oracle": so,
                "gpu_vs_oracle": d, "gpu_resort1_vs_resort7": d_self})
    assert np.isfinite(ref[0:6]).all()
    # the run moved the water: its statistics are far from the start's
    s0 = stats(parts)
    assert abs(s0["mean_y"] - so["mean_y"]) > 2.0
    assert d["mean_y"] < 0.25 and d["std_y"] < 0.25 and d["p05_y"] < 0.5 and d["p95_y"] < 0.5
    assert d["occupied"] < 0.03 and d["mean_per_cell"] < 0.03 and d["mean_speed"] < 0.05


# ---------------- dambreak128: re-sort every 10, whole run ----------------

def test_dambreak128_one_substep_frame_and_whole_run(gpu, oracle):
    sc = gpu.scenes.dambreak128()
    cfg, parts = sc["cfg"], sc["particles"]
    n = parts.shape[1]
    e = engine_for(gpu, sc, RESORT_EVERY=10)
    sdf = e.get_sdf()
    e.step()
    check_one_substep("dambreak128", one_substep_gaps(oracle, cfg, parts, sdf, e.getParticles()), n)
    e.close()

    # one frame (20 substeps, fused kernel and re-sorts inside) against the oracle
    k = int(cfg.SUBSTEPS)
    e = engine_for(gpu, sc, RESORT_EVERY=10)
    e.steps(k)
    got = e.getParticles()
    ref = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, k, nthreads=host_threads())
    gp, gv = max_err(got[0:3], ref[0:3]), max_err(got[3:6], ref[3:6])
    say(f"dambreak128: one frame of {k} substeps vs oracle: pos {gp:.3e} vel {gv:.3e}")
    parity_log({"workload": "dambreak128", "frame_substeps": k, "gpu_vs_fp32_frame": {"pos": gp, "vel": gv}})
    assert gp < k * POS_TOL and gv < k * VEL_TOL

    # the whole run: 200 frames at re-sort period 10, and at the reference's period 1
    finals = {}
    for resort, eng in ((10, e), (1, None)):
        if eng is None:
            eng = engine_for(gpu, sc, RESORT_EVERY=resort)
            eng.steps(k)
        t0 = time.time()
        for _ in range(sc["frames"] - 1):
            eng.steps(k)
        eng.synchronize()
        say(f"dambreak128 resort {resort}: {sc['frames']} frames, last {sc['frames'] - 1} in {time.time() - t0:.2f}s")
        finals[resort] = check_invariants(f"dambreak128 resort {resort}", eng, n,
                                          float(eng.getGrid()[0].astype(np.float64).sum()))
        eng.close()
    s10, s1 = stats(finals[10]), stats(finals[1])
    d = stat_gaps(s10, s1)
    s0 = stats(parts)
    say(f"dambreak128 after {sc['frames']} frames: resort 10 {s10}\n  resort 1 {s1}\n  gap {d}")
    parity_log({"workload": "dambreak128", "whole_run_substeps": sc["frames"] * k, "resort10": s10, "resort1": s1,
                "gap": d})
    assert abs(s0["mean_y"] - s1["mean_y"]) > 2.0  # the dam broke
    assert d["mean_y"] < 0.25 and d["std_y"] < 0.25 and d["p05_y"] < 0.5 and d["p95_y"] < 0.5
    assert d["occupied"] < 0.03 and d["mean_per_cell"] < 0.03 and d["mean_speed"] < 0.05


# ---------------- channel256: the roofline workload ----------------

def test_channel256_one_substep_and_one_frame(gpu, oracle):
    sc = gpu.scenes.channel256()
    cfg, parts = sc["cfg"], sc["particles"]
    n = parts.shape[1]
    assert n == 11_534_336
    e = engine_for(gpu, sc, RESORT_EVERY=sc["resort_every"])
    sdf = e.get_sdf()
    e.step()
    check_one_substep("channel256", one_substep_gaps(oracle, cfg, parts, sdf, e.getParticles()), n)
    e.close()

    k = int(cfg.SUBSTEPS)
    e = engine_for(gpu, sc, RESORT_EVERY=sc["resort_every"])
    e.steps(k)
    got = check_invariants("channel256 frame", e, n, float(e.getGrid()[0].astype(np.float64).sum()))
    e.close()
    t0 = time.time()
    ref = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float32), sdf, k, nthreads=host_threads())
    gp, gv = max_err(got[0:3], ref[0:3]), max_err(got[3:6], ref[3:6])
    say(f"channel256: one frame of {k} substeps vs oracle ({time.time() - t0:.1f}s): pos {gp:.3e} vel {gv:.3e}")
    parity_log({"workload": "channel256", "frame_substeps": k, "gpu_vs_fp32_frame": {"pos": gp, "vel": gv}})
    assert gp < k * POS_TOL and gv < k * VEL_TOL


# ---------------- batch64: 64 scenes, graph replay, autotuned CHUNK ----------------

def test_batch64_one_substep_per_scene_and_whole_run(gpu, oracle, tmp_path):
    from importlib import import_module

    at = import_module("water-sim-cuda_amd.autotune")
    b = gpu.scenes.batch_scenes()
    assert len(b["cfgs"]) == 64
    counts = b["counts"]
    e = gpu.Batch(b["cfgs"], RESORT_EVERY=b["resort_every"])
    e.init(counts, b["particles"])
    for s, sdf in enumerate(b["sdfs"]):
        e.set_scene_sdf(s, sdf)
    e.step()
    got = e.getParticles()
    e.close()
    worst = {"gpu_vs_fp32": {"pos": 0.0, "vel": 0.0}, "gpu_vs_fp64": {"pos": 0.0, "vel": 0.0},
             "fp32_vs_fp64": {"pos": 0.0, "vel": 0.0}}
    lo = 0
    for s, cfg in enumerate(b["cfgs"]):
        hi = lo + int(counts[s])
        g = one_substep_gaps(oracle, cfg, np.ascontiguousarray(b["particles"][:, lo:hi]), b["sdfs"][s],
                             got[:, lo:hi])
        for k1 in worst:
            for k2 in ("pos", "vel"):
                worst[k1][k2] = max(worst[k1][k2], g[k1][k2])
        lo = hi
    check_one_substep("batch64", worst, int(counts.sum()))

    e = gpu.Batch(b["cfgs"], RESORT_EVERY=b["resort_every"], GRAPH=1)
    e.init(counts, b["particles"])
    for s, sdf in enumerate(b["sdfs"]):
        e.set_scene_sdf(s, sdf)
    k = int(b["cfgs"][0].SUBSTEPS)
    e.steps(k)
    res = at.autotune(e, cache=tmp_path / "tune.json")
    say(f"batch64: autotune CHUNK {res['CHUNK']} LAUNCH_PCT {res['LAUNCH_PCT']}")
    t0 = time.time()
    for _ in range(b["frames"] - 1):
        e.steps(k)
    e.synchronize()
    say(f"batch64: {b['frames'] - 1} frames in {time.time() - t0:.2f}s, replays {e.get_option('GRAPH_REPLAYS')}")
    assert e.get_option("GRAPH_REPLAYS") >= b["frames"] - 1
    st = e.getParticles()
    chk = e.check(0.5)
    say(f"batch64: check {chk}")
    assert np.isfinite(st[0:24]).all() and chk["nonfinite"] == 0 and chk["penetrating"] == 0
    for s in range(64):
        info = e.scene_info(s)
        mass = float(e.scene_grid(s)[0].astype(np.float64).sum())
        assert abs(mass - info["count"]) < 1e-3 * info["count"], (s, mass, info)
    e.close()
