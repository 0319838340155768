"""Serving loop: continuous batching over concurrent slots (tdp1375_amd.serving).

CPU: the admission logic with the oracle standing in for the GPU launch
(requests that arrive during a launch are served by the next free slot, in one
batch).  GPU: every request's verdict is the same alone and in the mix, every
sequence replays, and a Poisson run reports throughput and latency.
"""
import threading
import time

import numpy as np
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import serving as S
from tdp1375_amd import workloads as W
from helpers import assert_witness


class FakeBackend:
    """search_batch on the oracle, taking `delay` seconds per launch."""

    def __init__(self, monkeypatch, delay):
        self.delay = delay
        self.batches = []
        self.lock = threading.Lock()
        monkeypatch.setattr(S.G, "init_slots", lambda *a: None)
        monkeypatch.setattr(S.G, "set_tuning", lambda *a, **k: None)
        monkeypatch.setattr(S.G, "destroy_slots", lambda: None)
        monkeypatch.setattr(S.G, "search_batch", self.search_batch)

    def search_batch(self, slot, pis, spis, e0s, rates, depths, full, flags=0):
        time.sleep(self.delay)
        res = [oracle.search(p, s, e, r, d) for p, s, e, r, d in zip(pis, spis, e0s, rates, depths)]
        with self.lock:
            self.batches.append((slot, len(pis)))
        return S.G.BatchResult(np.array([r.found for r in res]), np.array([r.depth for r in res]),
                             [r.moves for r in res],
                             {"expansions": np.array([r.counters["expansions"] for r in res], np.uint64)})


def test_requests_arriving_during_a_launch_share_the_next_one(monkeypatch):
    fake = FakeBackend(monkeypatch, delay=0.2)
    server = S.SearchServer(slots=1, max_batch=64)
    wl = W.workload_random_n12()
    reqs = [S.Request(c.pi, c.spi, c.even_cycles, wl.min_rate, wl.max_depth) for c in wl.cases[:17]]
    first = server.submit(reqs[0])
    time.sleep(0.05)                       # the first launch is running
    rest = [server.submit(r) for r in reqs[1:]]
    replies = [first.result()] + [f.result() for f in rest]
    server.close()
    assert [b for _, b in fake.batches] == [1, 16]
    want = [oracle.search(c.pi, c.spi, c.even_cycles, wl.min_rate, wl.max_depth).found for c in wl.cases[:17]]
    assert [r.found for r in replies] == want
    assert all(r.latency > 0 for r in replies) and replies[1].batch == 16


def test_slots_serve_concurrently(monkeypatch):
    fake = FakeBackend(monkeypatch, delay=0.3)
    server = S.SearchServer(slots=4, max_batch=1)
    wl = W.workload_random_n12()
    t0 = time.perf_counter()
    futs = [server.submit(S.Request(c.pi, c.spi, c.even_cycles, wl.min_rate, wl.max_depth)) for c in wl.cases[:4]]
    [f.result() for f in futs]
    elapsed = time.perf_counter() - t0
    server.close()
    assert sorted(s for s, _ in fake.batches) == [0, 1, 2, 3]
    assert elapsed < 0.3 * 3                 # four launches of 0.3 s overlapped


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix(gpu_lib):
    sat, unsat, rate = S.request_pool(2048, 5)
    assert len(sat) > 32 and len(unsat) > 32
    server = S.SearchServer(slots=8, queue_bytes_total=8 << 30, dedup_per_slot=1 << 22)
    try:
        pool = sat[:32] + unsat[:32]
        # alone: one launch per request on slot 0
        alone = []
        for c in pool:
            r = T.search_batch(0, [c.pi], [c.spi], [c.even_cycles], rate, 11, 0)
            alone.append(bool(r.found[0]))
        # in the mix: all at once through the server (batched, slots concurrent)
        futs = [server.submit(S.Request(c.pi, c.spi, c.even_cycles, rate, 11)) for c in pool]
        mixed = [f.result() for f in futs]
        assert [m.found for m in mixed] == alone == [True] * 32 + [False] * 32
        for c, m in zip(pool, mixed):
            if m.found:
                assert_witness(c.pi, c.spi, c.even_cycles, rate, 11, m.moves)
        assert max(m.batch for m in mixed) > 1
        out = S.poisson_run(server, pool, 200.0, 1.0, rate, 11, seed=1)
        print(out, flush=True)
        assert out["requests"] > 50 and out["p99_ms"] >= out["p50_ms"] > 0
    finally:
        server.close()
