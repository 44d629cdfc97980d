"""CPU tests of the online CHUNK tuner's schedule (no reference counterpart):
a stub engine whose frame time depends on CHUNK stands in for the device."""
from importlib import import_module

import pytest


class StubEngine:
    """steps(k) advances a fake clock by k times the per-substep cost of the
    current CHUNK."""

    def __init__(self, cost, chunk):
        self.cost, self.chunk = cost, chunk
        self.substeps, self.clock, self.sets = 0, 0.0, []

    def get_option(self, key):
        assert key == "CHUNK"
        return self.chunk

    def set_option(self, key, value):
        assert key == "CHUNK"
        self.chunk = value
        self.sets.append(value)

    def synchronize(self):
        pass

    def steps(self, k):
        self.substeps += k
        self.clock += self.cost[self.chunk] * k


@pytest.fixture()
def at(monkeypatch):
    return import_module("water-sim-cuda_amd.autotune")


def run(at, monkeypatch, cost, start, frames, period=2, chunks=(128, 256, 512, 1024)):
    e = StubEngine(cost, start)
    monkeypatch.setattr(at.time, "perf_counter", lambda: e.clock)
    t = at.OnlineTuner(e, 4, period=period, chunks=chunks)
    for _ in range(frames):
        t.frame()
    return e, t


def test_keeps_a_faster_neighbour_and_reverts_slower_ones(at, monkeypatch):
    cost = {128: 3.0, 256: 2.0, 512: 1.0, 1024: 1.5}
    e, t = run(at, monkeypatch, cost, 256, 8)
    # frame 1 tries 512 (kept), 3 tries 256, 5 tries 1024, 7 tries 256 (all reverted)
    assert e.substeps == 32 and t.frames == 8 and t.trials == 4 and t.switches == 1
    assert e.chunk == 512
    assert [h[1] for h in t.history] == [512, 256, 1024, 256]
    assert [h[4] for h in t.history] == [True, False, False, False]
    assert e.sets == [512, 256, 512, 1024, 512, 256, 512]


def test_margin_keeps_the_current_chunk_on_a_tie(at, monkeypatch):
    cost = {128: 1.0, 256: 0.99, 512: 1.0, 1024: 1.0}  # 1 % faster: inside the 3 % margin
    e, t = run(at, monkeypatch, cost, 512, 6)
    assert e.chunk == 512 and t.switches == 0 and t.trials == 3


def test_ends_of_the_list_try_the_only_neighbour(at, monkeypatch):
    cost = {128: 1.0, 256: 2.0}
    e, t = run(at, monkeypatch, cost, 128, 4, chunks=(128, 256))
    assert [h[1] for h in t.history] == [256, 256] and e.chunk == 128


def test_period_schedules_one_trial_per_period(at, monkeypatch):
    cost = {128: 1.0, 256: 1.0, 512: 1.0, 1024: 1.0}
    e, t = run(at, monkeypatch, cost, 256, 30, period=10)
    assert t.trials == 3 and e.substeps == 120
    assert [h[0] for h in t.history] == [9, 19, 29]


def test_rejects_empty_frames(at):
    with pytest.raises(ValueError):
        at.OnlineTuner(StubEngine({}, 256), 0)
