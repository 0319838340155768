"""Autotuner of the work-sharing launch (tdp1375_amd.autotune).

CPU: coordinate descent over a stand-in backend whose kernel time is a known
function of the knobs (the tuner must find its minimum along each knob, keep
the first upload out of the trials, refuse a trial that changes a verdict,
and round-trip the JSON cache).  GPU: a small grid on the n=12 batch keeps
every verdict, and the chosen tuning is the fastest trial.
"""
import json

import numpy as np
import pytest

from tdp1375_amd import autotune as A
from tdp1375_amd import workloads as W


class Fake:
    def __init__(self, monkeypatch, flip=None):
        self.knobs = {}
        self.calls = []
        self.flip = flip
        monkeypatch.setattr(A.G, "set_tuning", self.set_tuning)
        monkeypatch.setattr(A.G, "set_expansion_batch", lambda b=0: self.knobs.__setitem__("batch", b))
        monkeypatch.setattr(A.G, "search_batch", self.search_batch)
        monkeypatch.setattr(A.G, "last_launch_stats", lambda slot: {"kernel_ms": self.time()})

    def set_tuning(self, waves_per_cu=0, share_interval=0, share_max=0, idle_backoff=0, max_active=0):
        self.knobs.update(waves_per_cu=waves_per_cu, share_interval=share_interval, share_max=share_max,
                          max_active=max_active)

    def time(self):
        k = self.knobs
        return (100.0 + 10 * abs(k["share_max"] - 16) + 5 * abs(k["max_active"] - 512) / 512
                + (20 if k["batch"] != 8 else 0) + (3 if k["waves_per_cu"] else 0) + k["share_interval"] / 8)

    def search_batch(self, slot, pis, spis, e0s, rate, depth, full, flags=0):
        self.calls.append(flags)
        found = np.array([True, False, True])
        if self.flip and all(self.knobs.get(k) == v for k, v in self.flip.items()):
            found = np.array([True, True, True])
        return A.G.BatchResult(found, np.zeros(3, np.int32), [[], [], []], {"expansions": np.zeros(3, np.uint64)})


def test_coordinate_descent_finds_the_minimum_and_caches_it(monkeypatch, tmp_path):
    fake = Fake(monkeypatch)
    res = A.autotune(0, [[0]] * 3, [[0]] * 3, [0] * 3, 1.0, 4)
    assert res["best"] == {"waves_per_cu": 0, "share_interval": 8, "share_max": 16, "batch": 8, "max_active": 512}
    assert res["kernel_ms"] == min(t["kernel_ms"] for t in res["trials"])
    # the first run uploads the batch; every trial runs from device memory
    assert fake.calls[0] & A._lib.FLAG_RESIDENT == 0
    assert all(f & A._lib.FLAG_RESIDENT for f in fake.calls[1:])
    assert len(res["trials"]) == 1 + sum(len(v) - 1 for v in A.GRID.values())
    path = str(tmp_path / "tuning.json")
    A.save(path, "w", res)
    fake.knobs.clear()
    assert A.apply(path, "w") == res["best"]
    assert fake.knobs == res["best"]
    assert A.apply(path, "other") is None
    assert json.load(open(path))["w"]["best"] == res["best"]


class Coupled(Fake):
    """share_interval's best value depends on max_active, a knob swept after
    it: one pass settles share_interval too early."""

    def time(self):
        k = self.knobs
        want = 32 if k["max_active"] == 2048 else 8
        return 100.0 + abs(k["share_interval"] - want) + (0 if k["max_active"] == 2048 else 50)


def test_a_second_pass_revisits_knobs_after_a_later_one_moved(monkeypatch):
    Coupled(monkeypatch)
    one = A.autotune(0, [[0]] * 3, [[0]] * 3, [0] * 3, 1.0, 4, passes=1)
    assert one["best"]["max_active"] == 2048 and one["best"]["share_interval"] == 8
    two = A.autotune(0, [[0]] * 3, [[0]] * 3, [0] * 3, 1.0, 4, passes=3)
    assert two["best"]["max_active"] == 2048 and two["best"]["share_interval"] == 32
    assert two["kernel_ms"] == 100.0
    # no setting is measured twice
    seen = [tuple(sorted(t["knobs"].items())) for t in two["trials"]]
    assert len(seen) == len(set(seen))


def test_cache_keys_per_ordering():
    assert A.cache_key("w") == "w"
    assert A.cache_key("w", "learned") == "w/learned"


def test_a_trial_that_changes_a_verdict_is_an_error(monkeypatch):
    Fake(monkeypatch, flip={"share_max": 4})
    with pytest.raises(RuntimeError, match="verdicts changed"):
        A.autotune(0, [[0]] * 3, [[0]] * 3, [0] * 3, 1.0, 4)


@pytest.mark.gpu
def test_autotune_on_the_n12_batch_keeps_verdicts(gpu_lib):
    wl = W.workload_random_n12()
    A.G.init_slots(1, 1, 1 << 30, 1 << 20)
    try:
        # each knob: its default and one other value, so one trial per knob
        grid = {"share_max": [4, A.DEFAULT["share_max"]], "batch": [1, A.DEFAULT["batch"]],
                "max_active": [256, A.DEFAULT["max_active"]]}
        res = A.autotune(0, [c.pi for c in wl.cases], [c.spi for c in wl.cases], [c.even_cycles for c in wl.cases],
                         wl.min_rate, wl.max_depth, grid=grid)
        assert len(res["trials"]) == 1 + 3
        assert res["kernel_ms"] == min(t["kernel_ms"] for t in res["trials"])
    finally:
        A.reset()
        A.G.destroy_slots()
