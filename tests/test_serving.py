"""Serving through the device-resident request loop (tdp_serve_*,
tdp1375_amd.serving).

The request pool is the serving generator's configurations (n 9..48) that the
oracle decides within its budget, half satisfiable, with the oracle's labels
(tests/golden/serving_pool_n9-48-d11.json, written by make_serving_pool.py).
GPU: every reply equals its oracle label alone (one request in flight) and in
the mix (all at once), at a moderate and a saturating Poisson rate, with the
plain and the learned ordering; a cancelled request raises the reference's
error while the session keeps serving; concurrent searchForSortingSeq calls
from eight threads through the JNI natives share one session.  CPU: the pool
fixture and the Poisson harness.
"""
import json
import os
import threading
import time
from concurrent.futures import Future

import numpy as np
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import serving as S
from tdp1375_amd import workloads as W
from helpers import GOLDEN, assert_witness
from helpers import load_serving_pool as load_pool


def test_pool_fixture_is_half_satisfiable_and_matches_its_generator():
    sat, unsat, rate, doc = load_pool()
    assert len(sat) == len(unsat) == 128
    assert not set(doc["satisfiable"]) & set(doc["unsatisfiable"])
    for i in doc["satisfiable"] + doc["unsatisfiable"]:
        assert 9 <= doc["n"][str(i)] <= 48
    # relabel a few of the cheapest with the literal oracle
    for c, want in [(c, True) for c in sat[:6]] + [(c, False) for c in unsat[:6]]:
        r = oracle.search(c.pi, c.spi, c.even_cycles, rate, doc["max_depth"], max_expansions=2000)
        assert r.aborted or r.found == want


class _FakeServer:
    """submit() answers with the oracle at once (no GPU)."""

    def submit(self, req):
        f = Future()
        r = oracle.search(req.pi, req.spi, req.even_cycles, req.min_rate, req.max_depth)
        f.set_result(S.Reply(r.found, r.moves, r.counters["expansions"], time.perf_counter() - req.arrival))
        return f


def test_poisson_harness_maps_replies_to_requests():
    wl = W.workload_random_n12()
    pool = wl.cases[:20]
    out = S.poisson_run(_FakeServer(), pool, 400.0, 0.2, wl.min_rate, wl.max_depth, seed=3)
    assert out["requests"] == len(out["_replies"]) == len(out["_picks"]) > 10
    want = {k: oracle.search(c.pi, c.spi, c.even_cycles, wl.min_rate, wl.max_depth).found
            for k, c in enumerate(pool)}
    assert all(r.found == want[k] for r, k in zip(out["_replies"], out["_picks"]))
    assert out["p99_ms"] >= out["p50_ms"] >= 0


def _check_replies(pool, labels, replies, rate, max_depth):
    assert [r.found for r in replies] == labels
    for c, r in zip(pool, replies):
        if r.found:
            assert_witness(c.pi, c.spi, c.even_cycles, rate, max_depth, r.moves)


@pytest.mark.gpu
@pytest.mark.parametrize("learned", [False, True])
def test_serving_alone_mix_and_rates_match_oracle(gpu_lib, learned):
    sat, unsat, rate, doc = load_pool()
    D = doc["max_depth"]
    pool = sat + unsat
    labels = [True] * len(sat) + [False] * len(unsat)
    if learned:
        from tdp1375_amd import train_ordering
        train_ordering.load(os.path.join(os.path.dirname(S.__file__), "data", "ordering_mlp_d11.npz"))
    with S.DeviceServer(queue_bytes=16 << 30, dedup=1 << 26, slots=256, max_n=48, learned=learned,
                        lifetime_s=300) as srv:
        # alone: one request in flight at a time
        alone = [srv.search(c.pi, c.spi, c.even_cycles, rate, D) for c in pool[::4]]
        _check_replies(pool[::4], labels[::4], alone, rate, D)
        # in the mix: all at once
        futs = [srv.submit(S.Request(c.pi, c.spi, c.even_cycles, rate, D)) for c in pool]
        mixed = [f.result() for f in futs]
        _check_replies(pool, labels, mixed, rate, D)
        # Poisson arrivals at a moderate and at a saturating rate
        for offered, secs in ((200.0, 1.0), (5000.0, 0.5)):
            out = S.poisson_run(srv, pool, offered, secs, rate, D, seed=int(offered))
            assert [r.found for r in out["_replies"]] == [labels[k] for k in out["_picks"]]
            for r, k in zip(out["_replies"], out["_picks"]):
                if r.found:
                    assert_witness(pool[k].pi, pool[k].spi, pool[k].even_cycles, rate, D, r.moves)
            print(f"{'learned' if learned else 'plain'} offered {offered:.0f}/s: {out['requests']} requests, "
                  f"{out['throughput']:.0f}/s, p50 {out['p50_ms']:.2f} ms, p99 {out['p99_ms']:.2f} ms", flush=True)
        st = srv.stats()
        assert st["completed"] >= len(pool)


def _hard_request():
    """The depth-11 batch configuration the oracle left undecided with the
    most symbols (millions of expansions in the reference's order)."""
    with open(os.path.join(GOLDEN, "verdicts_odd-cycle-mix-n9-36-d11.json")) as fh:
        v = json.load(fh)["verdicts"]
    wl = W.workload_odd_cycle_mix_d11(4096)
    open_ = [c for c, x in zip(wl.cases, v) if x == "?"]
    return max(open_, key=lambda c: c.n), wl.min_rate


@pytest.mark.gpu
def test_serving_cancel_and_slot_routing(gpu_lib, monkeypatch):
    """With TDP_SERVE=1 initSlots starts a session and searches on the slots
    are served by it; cancelSort ends a running request with the reference's
    error and the session goes on serving."""
    monkeypatch.setenv("TDP_SERVE", "1")
    monkeypatch.setenv("TDP_SERVE_LIFETIME_S", "300")
    T.init_slots(2, 1, 4 << 30, 1 << 24)
    try:
        hard, rate = _hard_request()
        box = {}

        def run():
            try:
                box["res"] = T.search_for_sorting_seq(hard.pi, hard.spi, hard.even_cycles, rate, 11, False, 0)
            except Exception as e:  # noqa: BLE001
                box["err"] = e

        th = threading.Thread(target=run)
        th.start()
        time.sleep(0.02)
        T.cancel_sort(0)
        th.join(120)
        assert not th.is_alive()
        if "err" in box:
            assert isinstance(box["err"], T.CancelledError) and str(box["err"]) == "GPU sorting search cancelled"
        else:   # it finished before the cancel arrived: then its sequence must replay
            if box["res"].positions:
                assert_witness(hard.pi, hard.spi, hard.even_cycles, rate, 11, box["res"].positions)
        wl = W.workload_random_n12()
        for c in wl.cases[:16]:
            r = T.search_for_sorting_seq(c.pi, c.spi, c.even_cycles, wl.min_rate, 6, False, 1)
            assert bool(r.positions) == oracle.search(c.pi, c.spi, c.even_cycles, wl.min_rate, 6).found
    finally:
        T.destroy_slots()
