"""Autotuner for the engine's work decomposition (no reference counterpart).

It picks the transfer work-item size (``CHUNK``: particles per chunk, one
wave per chunk) and the transfer launch width (``LAUNCH_PCT``: workgroups
per 100 chunks) by timing a few substeps of the caller's own state for each
candidate, rewinding the state between trials (``Engine.snapshot`` /
``Engine.restore``), and leaves the engine on the fastest choice with its
state as it was before tuning.  Results are cached in a JSON file keyed by
the shape of the workload (scene sizes and particle counts), so later runs
of the same workload skip the search.
"""
from __future__ import annotations

import json
import math
import os
import time
from pathlib import Path

CHUNKS = (128, 256, 512, 1024, 2048)
LAUNCH_PCTS = (100, 50, 25)


def signature(eng) -> str:
    """Workload key: scene grid sizes and particle counts (rounded to 1/8
    octaves), in scene order."""
    cfgs = getattr(eng, "cfgs", [eng.cfg])
    if hasattr(eng, "scene_info"):
        counts = [eng.scene_info(s)["count"] for s in range(len(cfgs))]
    else:
        counts = [eng.getNumber()]
    parts = []
    for c, n in zip(cfgs, counts):
        q = 0 if n <= 0 else round(8 * math.log2(n))
        parts.append(f"{c.SIZE_X}x{c.SIZE_Y}x{c.SIZE_Z}:{q}")
    return "|".join(parts)


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def autotune(eng, trial_substeps: int = 8, warmup: int = 2, chunks=CHUNKS, launch_pcts=LAUNCH_PCTS,
             cache: str | os.PathLike | None = None) -> dict:
    """Tunes ``eng`` in place and returns {"CHUNK", "LAUNCH_PCT", "ms_per_substep", "trials", "cached"}."""
    key = signature(eng)
    path = Path(cache) if cache else None
    if path is not None:
        hit = _load(path).get(key)
        if hit:
            eng.set_option("CHUNK", hit["CHUNK"])
            eng.set_option("LAUNCH_PCT", hit["LAUNCH_PCT"])
            return dict(hit, cached=True)
    if eng.getNumber() == 0:
        return {"CHUNK": eng.get_option("CHUNK"), "LAUNCH_PCT": eng.get_option("LAUNCH_PCT"),
                "ms_per_substep": 0.0, "trials": [], "cached": False}
    eng.snapshot()
    trials = []
    for ch in chunks:
        for pct in launch_pcts:
            eng.set_option("CHUNK", ch)
            eng.set_option("LAUNCH_PCT", pct)
            eng.steps(warmup)
            eng.synchronize()
            t0 = time.perf_counter()
            eng.steps(trial_substeps)
            eng.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / trial_substeps
            trials.append({"CHUNK": ch, "LAUNCH_PCT": pct, "ms_per_substep": round(ms, 5)})
            eng.restore()
    best = min(trials, key=lambda t: t["ms_per_substep"])
    eng.set_option("CHUNK", best["CHUNK"])
    eng.set_option("LAUNCH_PCT", best["LAUNCH_PCT"])
    eng.restore()
    result = dict(best, trials=trials, cached=False)
    if path is not None:
        data = _load(path)
        data[key] = {k: best[k] for k in ("CHUNK", "LAUNCH_PCT", "ms_per_substep")}
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    return result
