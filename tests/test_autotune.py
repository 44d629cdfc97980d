"""CPU tests of the online CHUNK tuner's schedule (no reference counterpart):
a stub engine whose frame time depends on CHUNK stands in for the device."""
from importlib import import_module

import pytest


class StubEngine:
    """steps(k) advances a fake clock by k times the per-substep cost of the
    current CHUNK.  It models a one-entry graph cache: a steps() call whose
    CHUNK differs from the previous call's "captures" a graph."""

    def __init__(self, cost, chunk):
        self.cost, self.chunk = cost, chunk
        self.substeps, self.clock, self.sets = 0, 0.0, []
        self.captures, self._graph_chunk = 0, None

    def get_option(self, key):
        if key == "GRAPH_CAPTURES":
            return self.captures
        assert key == "CHUNK"
        return self.chunk

    def set_option(self, key, value):
        assert key == "CHUNK"
        self.chunk = value
        self.sets.append(value)

    def synchronize(self):
        pass

    def steps(self, k):
        if self._graph_chunk != self.chunk:
            self.captures += 1
            self._graph_chunk = self.chunk
        self.substeps += k
        self.clock += self.cost[self.chunk] * k


@pytest.fixture()
def at(monkeypatch):
    return import_module("water-sim-cuda_amd.autotune")


def run(at, monkeypatch, cost, start, frames, period=4, chunks=(128, 256, 512, 1024)):
    e = StubEngine(cost, start)
    monkeypatch.setattr(at.time, "perf_counter", lambda: e.clock)
    t = at.OnlineTuner(e, 4, period=period, chunks=chunks)
    for _ in range(frames):
        t.frame()
    return e, t


def test_keeps_a_faster_neighbour_and_reverts_slower_ones(at, monkeypatch):
    cost = {128: 3.0, 256: 2.0, 512: 1.0, 1024: 1.5}
    e, t = run(at, monkeypatch, cost, 256, 16)
    # frame 3 times 512 (kept), 7 times 256, 11 times 1024, 15 times 256 (all reverted)
    assert e.substeps == 64 and t.frames == 16 and t.trials == 4 and t.switches == 1
    assert e.chunk == 512
    assert [h[0] for h in t.history] == [3, 7, 11, 15]
    assert [h[1] for h in t.history] == [512, 256, 1024, 256]
    assert [h[4] for h in t.history] == [True, False, False, False]
    assert e.sets == [512, 256, 512, 1024, 512, 256, 512]


def test_timed_frames_never_include_a_capture(at, monkeypatch):
    """The frame after a CHUNK change (which captures the new graph) is not
    timed: neither timed frame of any trial captured."""
    cost = {128: 3.0, 256: 2.0, 512: 1.0, 1024: 1.5}
    e, t = run(at, monkeypatch, cost, 256, 40)
    assert t.trials == 10
    assert not any(h[5] for h in t.history)
    assert e.captures > t.trials  # every switch and every revert captured, untimed


class AlternatingStub(StubEngine):
    """A graph cache keyed by (CHUNK, call parity): the engine's start state
    alternates between calls, so a new CHUNK captures on two calls in a row;
    a CHUNK change drops every cached graph."""

    def __init__(self, cost, chunk):
        super().__init__(cost, chunk)
        self.cache, self.calls = set(), 0

    def set_option(self, key, value):
        super().set_option(key, value)
        self.cache.clear()

    def steps(self, k):
        key = (self.chunk, self.calls % 2)
        if key not in self.cache:
            self.captures += 1
            self.cache.add(key)
        self.calls += 1
        self.substeps += k
        self.clock += self.cost[self.chunk] * k


def test_timed_frame_after_a_capture_is_retried(at, monkeypatch):
    cost = {128: 3.0, 256: 2.0, 512: 1.0, 1024: 1.5}
    e = AlternatingStub(cost, 256)
    monkeypatch.setattr(at.time, "perf_counter", lambda: e.clock)
    t = at.OnlineTuner(e, 4, period=6, chunks=(128, 256, 512, 1024))
    for _ in range(30):
        t.frame()
    assert t.frames == 30 and t.trials >= 3
    # each candidate frame after the untimed switch frame captured the other
    # parity's graph once and was timed again on the next frame
    assert all(h[5] >= 1 for h in t.history)
    assert t.history[0][1] == 512 and t.history[0][4]
    assert e.chunk == 512


def test_margin_keeps_the_current_chunk_on_a_tie(at, monkeypatch):
    cost = {128: 1.0, 256: 0.99, 512: 1.0, 1024: 1.0}  # 1 % faster: inside the 3 % margin
    e, t = run(at, monkeypatch, cost, 512, 12)
    assert e.chunk == 512 and t.switches == 0 and t.trials == 3


def test_ends_of_the_list_try_the_only_neighbour(at, monkeypatch):
    cost = {128: 1.0, 256: 2.0}
    e, t = run(at, monkeypatch, cost, 128, 8, chunks=(128, 256))
    assert [h[1] for h in t.history] == [256, 256] and e.chunk == 128


def test_period_schedules_one_trial_per_period(at, monkeypatch):
    cost = {128: 1.0, 256: 1.0, 512: 1.0, 1024: 1.0}
    e, t = run(at, monkeypatch, cost, 256, 30, period=10)
    assert t.trials == 3 and e.substeps == 120
    assert [h[0] for h in t.history] == [9, 19, 29]


def test_short_periods_are_raised_to_four(at):
    assert at.OnlineTuner(StubEngine({256: 1.0}, 256), 4, period=2).period == 4


def test_rejects_empty_frames(at):
    with pytest.raises(ValueError):
        at.OnlineTuner(StubEngine({}, 256), 0)


class SnapEngine(StubEngine):
    """StubEngine with the snapshot/restore and LAUNCH_PCT the offline
    autotuner uses; restore() returns to the snapshot's CHUNK graph state."""

    def __init__(self, cost, chunk):
        super().__init__(cost, chunk)
        self.pct = 100
        self.cfg = type("Cfg", (), {"SIZE_X": 32, "SIZE_Y": 32, "SIZE_Z": 32})()

    def get_option(self, key):
        if key == "LAUNCH_PCT":
            return self.pct
        return super().get_option(key)

    def set_option(self, key, value):
        if key == "LAUNCH_PCT":
            self.pct = value
            return
        super().set_option(key, value)

    def getNumber(self):
        return 1000

    def snapshot(self):
        pass

    def restore(self):
        pass


def test_autotune_trials_time_replays_not_captures(at, monkeypatch):
    cost = {128: 3.0, 256: 2.0, 512: 1.0, 1024: 1.5, 2048: 1.2}
    e = SnapEngine(cost, 2048)
    monkeypatch.setattr(at.time, "perf_counter", lambda: e.clock)
    r = at.autotune(e, trial_substeps=8, warmup=2, launch_pcts=(100,))
    assert r["CHUNK"] == 512 and not r["cached"]
    assert len(r["trials"]) == 5
    assert not any(t["captured"] for t in r["trials"])
    assert [t["ms_per_substep"] for t in r["trials"]] == [3.0e3, 2.0e3, 1.0e3, 1.5e3, 1.2e3]
