"""Autotuner for the engine's work decomposition (no reference counterpart).

It picks the transfer work-item size (``CHUNK``: particles per chunk, one
wave per chunk) and the transfer launch width (``LAUNCH_PCT``: workgroups
per 100 chunks) by timing a few substeps of the caller's own state for each
candidate, rewinding the state between trials (``Engine.snapshot`` /
``Engine.restore``), and leaves the engine on the fastest choice with its
state as it was before tuning.  Results are cached in a JSON file keyed by
the shape of the workload (scene sizes and particle counts), so later runs
of the same workload skip the search.

The fastest CHUNK can change as a run goes on: a chunk needs as many rounds
as its fullest base cell holds particles, so where water compresses into
dense cells, splitting blocks into more, smaller chunks gives the GPU more
waves at no cost in rounds, while in the fresh state whole blocks win.
``OnlineTuner`` therefore keeps tuning CHUNK during the run itself.
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


def _captures(eng) -> int:
    """Graphs the engine has captured so far (0 for engines without graphs)."""
    try:
        return int(eng.get_option("GRAPH_CAPTURES"))
    except Exception:
        return 0


def _settle(eng, k: int, tries: int = 3) -> None:
    """Runs untimed steps(k) calls until one of them replays a cached graph
    instead of capturing a new one (at most `tries` calls; one call without
    graph replay), so that a following timed steps(k) call measures a replay
    and not a capture."""
    for _ in range(tries):
        before = _captures(eng)
        eng.steps(k)
        if _captures(eng) == before:
            break


def autotune(eng, trial_substeps: int = 8, warmup: int = 2, chunks=CHUNKS, launch_pcts=LAUNCH_PCTS,
             cache: str | os.PathLike | None = None) -> dict:
    """Tunes ``eng`` in place and returns {"CHUNK", "LAUNCH_PCT", "ms_per_substep", "trials", "cached"}.

    Each trial sets the candidate, runs untimed steps(trial_substeps) calls
    (at least ``warmup`` substeps, and with graph replay until a call replays
    its graph), then times one steps(trial_substeps) call; a trial whose
    timed call still captured a graph is marked ``captured``."""
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
            for _ in range(max(1, -(-warmup // trial_substeps))):
                eng.steps(trial_substeps)
            _settle(eng, trial_substeps)
            eng.synchronize()
            c0 = _captures(eng)
            t0 = time.perf_counter()
            eng.steps(trial_substeps)
            eng.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / trial_substeps
            trials.append({"CHUNK": ch, "LAUNCH_PCT": pct, "ms_per_substep": round(ms, 5),
                           "captured": _captures(eng) != c0})
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


class OnlineTuner:
    """Runs a caller's frames (``frame_substeps`` substeps each) and tunes
    CHUNK on the way.  Once per ``period`` frames (at least 4) a trial runs
    over consecutive frames: one frame is timed at the current CHUNK; the
    next moves CHUNK to a neighbouring value in ``chunks`` (alternately the
    next larger and the next smaller) and runs untimed, taking the change's
    re-sort, the new layout's first launches and, with graph replay, the
    capture of the new frame graph; the frame after it is timed at the
    neighbour, which is kept when it was faster by more than ``margin``,
    otherwise CHUNK goes back.  A timed frame that captured a graph (the
    engine's start state can alternate between calls, so a new setting may
    need one capture per state) does not count: the next frame is timed in
    its place, up to ``retries`` times.  Nothing is rewound: every frame is a
    frame of the run, and only the timed frames synchronise with the device.
    With no graph replay the trial occupies the last three frames of each
    period."""

    def __init__(self, eng, frame_substeps: int, period: int = 10, margin: float = 0.03, chunks=CHUNKS,
                 retries: int = 3):
        if frame_substeps < 1:
            raise ValueError("frame_substeps must be at least 1")
        self.eng = eng
        self.k = int(frame_substeps)
        self.period = max(4, int(period))
        self.margin = float(margin)
        self.retries = int(retries)
        self.chunks = tuple(sorted(set(int(c) for c in chunks)))
        self.frames = 0
        self.trials = 0
        self.switches = 0
        # (frame, CHUNK tried, its ms, the current CHUNK's ms, kept, timed frames retried for a capture)
        self.history = []
        self._up = True
        self._idle = 0           # frames since the last trial ended
        self._state = "idle"
        self._tries = 0
        self._retried = 0
        self._ref_ms = None
        self._cur = None
        self._cand = None

    def _neighbour(self, cur: int):
        lower = [c for c in self.chunks if c < cur]
        higher = [c for c in self.chunks if c > cur]
        first, second = (higher, lower) if self._up else (lower, higher)
        self._up = not self._up
        if first:
            return first[0] if first is higher else first[-1]
        if second:
            return second[0] if second is higher else second[-1]
        return None

    def _timed(self):
        e = self.eng
        e.synchronize()
        c0 = _captures(e)
        t0 = time.perf_counter()
        e.steps(self.k)
        e.synchronize()
        return (time.perf_counter() - t0) * 1e3, _captures(e) != c0

    def frame(self) -> None:
        """Runs one frame."""
        e = self.eng
        self.frames += 1
        if self._state == "idle":
            self._idle += 1
            if self._idle >= self.period - 3:
                self._state, self._tries = "ref", 0
            e.steps(self.k)
            return
        if self._state == "ref":
            ms, cap = self._timed()
            if cap and self._tries < self.retries:
                self._tries += 1
                self._retried += 1
                return
            self._ref_ms = ms
            self._cur = e.get_option("CHUNK")
            self._cand = self._neighbour(self._cur)
            self._state = "switch" if self._cand is not None else "idle"
            if self._cand is None:
                self._idle = 0
            return
        if self._state == "switch":
            e.set_option("CHUNK", self._cand)
            e.steps(self.k)
            self._state, self._tries = "cand", 0
            return
        # "cand"
        ms, cap = self._timed()
        if cap and self._tries < self.retries:
            self._tries += 1
            self._retried += 1
            return
        self.trials += 1
        kept = ms < self._ref_ms * (1.0 - self.margin)
        if kept:
            self.switches += 1
        else:
            e.set_option("CHUNK", self._cur)
        self.history.append((self.frames - 1, self._cand, round(ms, 4), round(self._ref_ms, 4), kept,
                             self._retried))
        self._state, self._idle, self._retried = "idle", 0, 0
        self._ref_ms = self._cand = None
