"""Benchmark: MLS-MPM substeps of the 256^3 channel workload on one MI355X.

Prints one JSON line.  A "step" is one MLS-MPM substep (P2G, grid update,
G2P, and the Morton re-sort every RESORT_EVERY substeps) over all particles,
with the state resident in HBM.  The run is the reference's frame loop:
substeps are issued SUBSTEPS at a time per engine call (40 for the 256^3
channel), so within a call the fused G2P+P2G kernel runs once per substep
boundary.

* `value` is particle-substeps per second over exactly `--steps` timed
  substeps after exactly `--warmup` untimed ones (the last call of each part
  is shortened when the count is not a multiple of SUBSTEPS).
* When `--steps` is shorter than the workload's whole run (frames x SUBSTEPS,
  100 x 40 for the 256^3 channel), the JSON line also carries `sustained`:
  the whole run timed the same way from a re-initialised state after two
  warm-up frames, with its own roofline.  With no flags `value` is that run.
* `roofline` describes the dominant kernel (fused G2P+P2G), timed with HIP
  events on the engine's stream inside the timed region.
* `cpu_baseline` times the oracle's OpenMP step on the host threads this
  process may use, on a bounded sample (a few substeps of the same scene).

`--scene batch64` measures the batched 64-scene workload instead, with its
block and tile sizes chosen by the autotuner as that workload specifies
(trials on the warmed-up state, then a CHUNK trial every 10 frames of the
run; `--no-tune` keeps the hand-set CHUNK); `--scene dambreak128` is tuned
the same way; `--graph` replays each frame as a HIP graph; `--autotune` and
`--tune-online F` tune any scene.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-substeps/s and % HBM roofline at 1 GPU; learned-4xDT vs plain sim-s/s at RMS err"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per launch of the fused G2P+P2G kernel: per particle it
# reads position (12 B) and F (36 B) and writes both back (48 B); v and C
# stay in registers.  Per active grid node it reads the velocity (12 B) and
# the SDF (4 B) and writes mass and momentum (16 B).
BYTES_PER_PARTICLE = 96
BYTES_PER_ACTIVE_NODE = 32


def read_traffic(path: str, workload: str):
    """HBM bytes per fused-kernel launch from a profile summary ('key: value'
    lines), if it was measured on this workload."""
    p = Path(path)
    if not p.exists():
        return None
    kv = {}
    for line in p.read_text().splitlines():
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    if kv.get("workload") != workload or "hbm_bytes_per_launch" not in kv:
        return None
    return int(float(kv["hbm_bytes_per_launch"]))


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> int:
    """Host CPUs this process may use: its affinity set, capped by the
    cgroup CPU quota when one is set."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_substeps(eng, total: int, frame: int, tuner=None) -> None:
    """`total` substeps in calls of `frame` (the last one shorter); with a
    tuner, whole frames go through it."""
    done = 0
    while done < total:
        k = min(frame, total - done)
        if tuner is not None and k == frame:
            tuner.frame()
        else:
            eng.steps(k)
        done += k


def timed_run(eng, steps: int, frame: int, tuner=None) -> tuple[float, dict]:
    """Times exactly `steps` substeps; returns (seconds, stage events)."""
    eng.synchronize()
    eng.stage_times()  # drop earlier events
    t0 = time.perf_counter()
    run_substeps(eng, steps, frame, tuner)
    eng.synchronize()
    return time.perf_counter() - t0, eng.stage_times()


def roofline(st: dict, n: int, active: int, traffic) -> dict | None:
    ms, cnt = st["g2p2g"]
    if cnt <= 0:
        return None
    avg_s = ms / cnt / 1e3
    bytes_launch = BYTES_PER_PARTICLE * n + BYTES_PER_ACTIVE_NODE * active
    achieved = bytes_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_transfer<G2P2G>", "launches": cnt, "avg_launch_ms": round(ms / cnt, 5),
            "algorithmic_bytes_per_launch": bytes_launch, "active_nodes": active}


def active_nodes(eng) -> int:
    """Grid nodes with mass > 1e-9 after the last grid update (the A of the
    algorithmic byte count)."""
    return int((eng.getGrid()[0] > 1e-9).sum())


def cpu_baseline(sc: dict, batch: bool, sdf, seconds: float) -> dict:
    """The oracle's OpenMP step (the reference algorithm, restated on the
    host) timed on this process's host threads, on a bounded sample of the
    same workload."""
    from oracle import oracle

    nthreads = host_threads()
    omp_max = int(oracle.lib().oracle_max_threads())
    model = cpu_model()
    if batch:
        # scene by scene (the oracle steps one scene at a time) until the
        # time budget is used; 2 substeps per scene
        done, t_sum, used, lo = 0, 0.0, 0, 0
        for s_, c_ in enumerate(sc["cfgs"]):
            hi = lo + int(sc["counts"][s_])
            st0 = oracle.init_state(c_, np.ascontiguousarray(sc["particles"][:, lo:hi]), np.float32)
            t_k, _ = oracle.time_steps(c_, st0, sc["sdfs"][s_], 2, nthreads)
            done += 2 * (hi - lo)
            t_sum += t_k
            used += 1
            lo = hi
            if t_sum > seconds:
                break
        value, sample = done / t_sum, (f"2 substeps of each of the first {used} of {len(sc['cfgs'])} scenes "
                                       f"({done // 2} particles) with the OpenMP oracle, {t_sum:.1f}s wall")
    else:
        cfg, n = sc["cfg"], sc["particles"].shape[1]
        st0 = oracle.init_state(cfg, sc["particles"], np.float32)
        t_one, _ = oracle.time_steps(cfg, st0, sdf, 1, nthreads)
        k = max(1, min(20, int(seconds / max(t_one, 1e-3))))
        t_k, _ = oracle.time_steps(cfg, st0, sdf, k, nthreads)
        value, sample = n * k / t_k, (f"{k} substeps of the full {sc['name']} scene (n={n}) with the OpenMP "
                                      f"oracle, {t_k:.1f}s wall")
    return {"value": round(value, 1), "unit": "particle-substeps/s", "cores": nthreads, "kind": "port",
            "sample": sample, "cpu_model": model, "omp_max_threads": omp_max, "threads_used": nthreads}


def make_engine(mpm, sc: dict, batch: bool, opts: dict):
    if batch:
        eng = mpm.Batch(sc["cfgs"], **opts)
        eng.init(sc["counts"], sc["particles"])
        for s_, sdf in enumerate(sc["sdfs"]):
            eng.set_scene_sdf(s_, sdf)
    else:
        eng = mpm.Engine(sc["cfg"], **opts)
        eng.init(sc["particles"].shape[1], sc["particles"])
        if sc.get("sdf") is not None:
            eng.set_sdf(sc["sdf"])
        else:
            eng.initSDF(sc["triangles"])
    return eng


def reinit(eng, sc: dict, batch: bool) -> None:
    """Back to the scene's initial particle state (the SDF stays)."""
    if batch:
        eng.init(sc["counts"], sc["particles"])
    else:
        eng.init(sc["particles"].shape[1], sc["particles"])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed substeps [the workload's whole run: frames x SUBSTEPS]")
    ap.add_argument("--warmup", type=int, default=None, help="untimed substeps first [two frames]")
    ap.add_argument("--scene", default="channel256")
    ap.add_argument("--resort-every", type=int, default=None, help="re-sort period [the scene's]")
    ap.add_argument("--chunk", type=int, default=None, help="CHUNK [the scene's, else the engine default]")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events")
    ap.add_argument("--no-sustained", action="store_true", help="skip the whole-run measurement")
    ap.add_argument("--no-stage-table", action="store_true", help="skip the per-stage timing pass")
    ap.add_argument("--graph", action="store_true", help="replay each frame's steps() call as a HIP graph")
    ap.add_argument("--frame", type=int, default=0, help="substeps per steps() call [the scene's SUBSTEPS]")
    ap.add_argument("--autotune", action="store_true",
                    help="pick CHUNK and LAUNCH_PCT by timing trials first [on for the batch and dam-break scenes]")
    ap.add_argument("--no-tune", action="store_true",
                    help="no autotuning of the batch and dam-break scenes (their hand-set CHUNK)")
    ap.add_argument("--tune-cache", default=None, help="JSON cache file for --autotune")
    ap.add_argument("--tune-online", type=int, default=None, metavar="FRAMES",
                    help="keep tuning CHUNK in the timed run: one trial every FRAMES frames (0: off) "
                         "[10 for the batch and dam-break scenes, else 0]")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "g2p2g_channel256_traffic.txt"),
                    help="rocprofv3-derived HBM bytes per launch of the fused kernel (key: value text)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-corrector evaluation")
    ap.add_argument("--learned-seconds", type=float, default=1.0, help="simulated seconds of the learned evaluation")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the calibration copy of BYTES and print its bandwidth")
    ap.add_argument("--copy-wide", action="store_true",
                    help="with --copy-probe: the wide copy (4 loads in flight per thread, nontemporal stores)")
    args = ap.parse_args()

    if args.copy_probe:
        mpm = importlib.import_module("water-sim-cuda_amd")
        ms = mpm.probe_copy(args.copy_probe, 20, wide=args.copy_wide)
        print(json.dumps({"copy_bytes": args.copy_probe, "kernel": "k_copy_wide" if args.copy_wide else "k_copy",
                          "ms_per_copy": ms, "GBps_read_plus_write": 2 * args.copy_probe / ms / 1e6}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the engine is single-GPU (a multi-GPU slab decomposition "
              "is designed but not implemented); use --gpus 1",
              file=sys.stderr)
        return 2

    mpm = importlib.import_module("water-sim-cuda_amd")
    t0 = time.time()
    batch = args.scene.startswith("batch")
    if batch:
        sc = mpm.scenes.batch_scenes(count=int(args.scene[5:] or 64))
        cfg = sc["cfgs"][0]
        n = int(sc["counts"].sum())
        grid_desc = sorted({c.SIZE_X for c in sc["cfgs"]})
    else:
        sc = mpm.scenes.SCENES[args.scene]()
        cfg = sc["cfg"]
        n = sc["particles"].shape[1]
        grid_desc = [cfg.SIZE_X, cfg.SIZE_Y, cfg.SIZE_Z]
    resort = sc["resort_every"] if args.resort_every is None else args.resort_every
    # the batched workload is defined with block and tile sizes chosen by the
    # autotuner, and the dam break's splash wants other chunks than its fresh
    # block: such scenes are tuned first on the warmed-up state, then online
    # during the run
    if sc.get("tune") and not args.no_tune:
        args.autotune = True
        if args.tune_online is None:
            args.tune_online = 10
    if args.tune_online is None or args.no_tune:
        args.tune_online = 0
    log(f"scene {sc['name']}: n={n} grid={grid_desc} (host gen {time.time() - t0:.1f}s)")

    # The reference's frame loop: SUBSTEPS substeps per call (Simulation.cpp:42-54).
    frame = args.frame or int(cfg.SUBSTEPS) or 20
    run_total = int(sc["frames"]) * frame  # the workload's whole run
    steps = run_total if args.steps is None else max(1, args.steps)
    warmup = 2 * frame if args.warmup is None else max(0, args.warmup)

    # TIMING 2: HIP events around the transfer launches only (the roofline's
    # kernel); the per-stage table comes from a short TIMING 1 pass afterwards
    opts = {"RESORT_EVERY": resort, "TIMING": 0 if args.no_timing else 2, "GRAPH": 1 if args.graph else 0}
    if args.chunk or sc.get("chunk"):
        opts["CHUNK"] = args.chunk or sc["chunk"]
    t0 = time.time()
    eng = make_engine(mpm, sc, batch, opts)
    log(f"engine ready ({time.time() - t0:.2f}s)")
    traffic = read_traffic(args.traffic, sc["name"])

    run_substeps(eng, warmup, frame)
    eng.synchronize()
    tuned = None
    if args.autotune:
        # trials on the warmed-up state (each restored afterwards); the
        # chosen setting then runs the timed substeps
        tuned = importlib.import_module("water-sim-cuda_amd.autotune").autotune(eng, cache=args.tune_cache)
        log(f"autotune: CHUNK={tuned['CHUNK']} LAUNCH_PCT={tuned['LAUNCH_PCT']} cached={tuned['cached']} "
            f"trials {tuned.get('trials')}")
    log(f"warmup {warmup} substeps done")

    def make_tuner():
        if args.tune_online <= 0:
            return None
        return importlib.import_module("water-sim-cuda_amd.autotune").OnlineTuner(eng, frame,
                                                                                 period=args.tune_online)

    tuner = make_tuner()
    elapsed, st = timed_run(eng, steps, frame, tuner)
    if tuner is not None:
        log(f"online tuning: {tuner.trials} trials, {tuner.switches} switches, history {tuner.history}")
    ctr = eng.counters()
    active = active_nodes(eng)
    state = eng.getParticles()
    finite = bool(np.isfinite(state[0:6]).all())
    log(f"timed {steps} substeps in {elapsed:.3f}s; stages {st}; counters {ctr}; active nodes {active}")
    value = n * steps / elapsed
    roof = roofline(st, n, active, traffic) if not args.no_timing else None
    online = None if tuner is None else {"period_frames": tuner.period, "trials": tuner.trials,
                                         "switches": tuner.switches}

    sustained = None
    if steps < run_total and not args.no_sustained:
        # the workload's whole run, from the initial state, after two frames
        reinit(eng, sc, batch)
        if tuned is not None:
            eng.set_option("CHUNK", tuned["CHUNK"])
            eng.set_option("LAUNCH_PCT", tuned["LAUNCH_PCT"])
        run_substeps(eng, 2 * frame, frame)
        tuner_s = make_tuner()
        el_s, st_s = timed_run(eng, run_total, frame, tuner_s)
        act_s = active_nodes(eng)
        sustained = {"value": round(n * run_total / el_s, 1), "steps": run_total, "warmup": 2 * frame,
                     "ms_per_step": round(el_s / run_total * 1e3, 5),
                     "frames": int(sc["frames"]), "substeps_per_frame": frame,
                     "roofline": roofline(st_s, n, act_s, traffic) if not args.no_timing else None,
                     "state_finite": bool(np.isfinite(eng.getParticles()[0:6]).all())}
        log(f"sustained: {sustained}")

    chunk_used = eng.get_option("CHUNK")
    stage_tab = None
    if not args.no_timing and not args.no_stage_table:
        # every stage timed, after the timed regions (events between all
        # launches cost a few percent of the substep rate)
        eng.set_option("TIMING", 1)
        eng.stage_times()
        run_substeps(eng, max(frame, 240 // frame * frame), frame)
        eng.synchronize()
        stage_tab = {k: (round(v[0] / v[1], 5) if v[1] else None) for k, v in eng.stage_times().items()}
        log(f"stage table (TIMING 1 pass): {stage_tab}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc, batch, None if batch else eng.get_sdf(), args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
    eng.close()

    learned = None
    if not args.no_learned and not batch:
        # The learned 4x-DT form against the plain step on a held-out member of
        # the 128^3 dam-break family (weights trained by train_corrector.py).
        L = importlib.import_module("water-sim-cuda_amd.learned")
        wpath = Path(mpm.__file__).parent / "corrector_dambreak128.npz"
        corr, meta = L.load(wpath)
        ev = L.evaluate(corr, 1000, sim_seconds=args.learned_seconds)
        corr.close()
        learned = {"workload": f"dambreak128 held-out seed 1000, {args.learned_seconds:g} simulated second(s)",
                   "ratio": ev["ratio"], "plain_dt": ev["dt"],
                   "plain_sim_s_per_s": round(ev["plain_sim_s_per_s"], 4),
                   "learned_sim_s_per_s": round(ev["learned_sim_s_per_s"], 4),
                   "speedup": round(ev["learned_speedup"], 4),
                   "learned_rms_cells": round(ev["learned_rms"], 5),
                   "plain_coarse_sim_s_per_s": round(ev["plain_coarse_sim_s_per_s"], 4),
                   "plain_coarse_rms_cells": round(ev["plain_coarse_rms"], 5),
                   "noise_floor_rms_cells": ev["noise_floor_rms"],
                   "fallback_frames": ev["fallback_frames"], "frames": ev["frames"],
                   "training": {k: meta.get(k) for k in ("train_seeds", "samples", "epochs", "loss_last", "init")}}
        log(f"learned: {learned}")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc["name"], "grid": grid_desc, "particles": n,
                   "scenes": len(sc["cfgs"]) if batch else 1, "dt": cfg.DT, "resort_every": resort,
                   "chunk": chunk_used, "substeps_per_call": frame, "whole_run_substeps": run_total,
                   "graph": bool(args.graph), "autotune": None if tuned is None else
                   {k: tuned[k] for k in ("CHUNK", "LAUNCH_PCT", "cached")},
                   "online_tune": online,
                   "description": sc["description"], "state_finite": finite},
        "roofline": roof,
        "sustained": sustained,
        "cpu_baseline": cpu,
        "learned": learned,
        "stage_ms_per_launch": stage_tab,
    }
    if sustained is not None:
        out["note"] = (f"value times the requested {steps} substeps after {warmup}; 'sustained' is the "
                       f"workload's whole run ({run_total} substeps) timed the same way")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
