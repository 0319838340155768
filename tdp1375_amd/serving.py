"""Serving loop: single-configuration search requests over concurrent slots.

The reference serves the proof generator's ForkJoinPool: every worker thread
takes a slot from a pool, runs one search on it and gives it back
(GPUSortingSearch.java:19-28,47-87).  Here the same slots run continuous
batches:

* requests go into one admission queue;
* every slot has a worker thread; a worker whose slot is free takes every
  request waiting at that moment (up to max_batch) and runs them as one
  persistent launch (tdp_search_batch), so requests that arrive while other
  launches run are admitted as soon as a slot frees up;
* the slots' launches run concurrently on the GPU on their own streams, each
  with a share of the compute units (waves per CU / slots) and of the queue
  memory (the queue budget is split evenly over the slots at init), and every
  slot keeps its device buffers for the whole session.

A request's verdict does not depend on what it is batched with: batching only
changes which wave takes which node first.  Latency is measured from arrival
to the end of the launch that served the request.
"""
from __future__ import annotations

import argparse
import json
import queue
import sys
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from . import gpu_sorting_search as G
from . import workloads as W


@dataclass
class Request:
    pi: Sequence[int]
    spi: Sequence[int]
    even_cycles: int
    min_rate: float
    max_depth: int
    arrival: float = 0.0
    future: Future = field(default_factory=Future)


@dataclass
class Reply:
    found: bool
    moves: list            # positions (i, j, k) per move
    expansions: int
    latency: float         # seconds from arrival to completion
    batch: int             # size of the launch that served it
    slot: int


class SearchServer:
    def __init__(self, slots: int = 8, queue_bytes_total: int = 8 << 30, dedup_per_slot: int = 1 << 22,
                 max_batch: int = 256, learned: bool = False, waves_per_cu: int = 0):
        self.slots = slots
        self.max_batch = max_batch
        self.flags = _lib.FLAG_LEARNED if learned else 0
        G.init_slots(slots, 1, queue_bytes_total // slots, dedup_per_slot)
        # a share of the compute units per slot so that the launches overlap;
        # small launches of single requests do better with hand-offs of 4
        # nodes than with the default 8 (measured: 4,721 against 3,450 req/s
        # at saturation)
        G.set_tuning(waves_per_cu=waves_per_cu or max(1, 12 // slots), share_max=4)
        self._q: "queue.Queue[Optional[Request]]" = queue.Queue()
        self._stop = False
        self.launches = 0
        self._lock = threading.Lock()
        self._workers = [threading.Thread(target=self._work, args=(s,), daemon=True) for s in range(slots)]
        for w in self._workers:
            w.start()

    def submit(self, req: Request) -> Future:
        if not req.arrival:
            req.arrival = time.perf_counter()
        self._q.put(req)
        return req.future

    def _work(self, slot: int) -> None:
        while True:
            first = self._q.get()
            if first is None:
                self._q.put(None)   # the stop mark stays for the other workers
                return
            batch = [first]
            while len(batch) < self.max_batch:
                try:
                    r = self._q.get_nowait()
                except queue.Empty:
                    break
                if r is None:
                    self._q.put(None)   # leave the stop mark for the other workers
                    break
                batch.append(r)
            try:
                res = G.search_batch(slot, [r.pi for r in batch], [r.spi for r in batch],
                                     [r.even_cycles for r in batch], [r.min_rate for r in batch],
                                     [r.max_depth for r in batch], 0, flags=self.flags)
                done = time.perf_counter()
                with self._lock:
                    self.launches += 1
                for c, r in enumerate(batch):
                    r.future.set_result(Reply(bool(res.found[c]), res.moves[c], int(res.counters["expansions"][c]),
                                              done - r.arrival, len(batch), slot))
            except BaseException as exc:   # every request of a failed launch gets the error
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(exc)

    def close(self) -> None:
        self._q.put(None)
        for w in self._workers:
            w.join()
        G.set_tuning()
        G.destroy_slots()


def request_pool(count: int, seed: int, max_depth: int = 11, max_ratio: float = 1.375,
                 max_expansions: int = 200000):
    """Configurations with n uniform in 9..48 (the depth-11 generator's odd
    cycles), split by the plain search's verdict into satisfiable and
    unsatisfiable ones by a batched plain search in a session of its own (call
    it before a SearchServer exists).  Configurations that took more than
    max_expansions to decide are left out, so that a request costs
    milliseconds rather than seconds."""
    cases = W.odd_cycle_mix(count, seed, 9, 48)
    rate = W.Workload("pool", cases, max_ratio, max_depth, 0, 0, 1).min_rate
    G.init_slots(1, 1, 16 << 30, 1 << 26)
    try:
        # configurations still open after 2 s of a launch are left out
        G.set_launch_deadline(2000)
        res = G.search_batch(0, [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases],
                             rate, max_depth, 0, flags=_lib.FLAG_SOFT_DEADLINE)
    finally:
        G.destroy_slots()
    light = (res.counters["expansions"] <= max_expansions) & (res.depth >= 0)
    sat = [c for c, f, ok in zip(cases, res.found, light) if f and ok]
    unsat = [c for c, f, ok in zip(cases, res.found, light) if not f and ok]
    return sat, unsat, rate


def poisson_run(server: SearchServer, pool: List, rate_per_s: float, duration_s: float, min_rate: float,
                max_depth: int, seed: int) -> dict:
    """Poisson arrivals for duration_s; returns throughput and latency percentiles."""
    rng = np.random.default_rng(seed)
    futures = []
    t0 = time.perf_counter()
    t_next = t0
    while True:
        t_next += rng.exponential(1.0 / rate_per_s)
        if t_next - t0 > duration_s:
            break
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        c = pool[int(rng.integers(len(pool)))]
        futures.append(server.submit(Request(c.pi, c.spi, c.even_cycles, min_rate, max_depth, arrival=t_next)))
    replies = [f.result() for f in futures]
    wall = time.perf_counter() - t0
    lat = np.array([r.latency for r in replies]) * 1e3
    return {"offered_rate": rate_per_s, "requests": len(replies), "throughput": len(replies) / wall,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_batch": float(np.mean([r.batch for r in replies])) if replies else None,
            "found": int(sum(r.found for r in replies))}


def main(argv=None):
    ap = argparse.ArgumentParser(description="Poisson serving sweep over concurrent slots")
    ap.add_argument("--rates", default="10,100,1000,10000")
    ap.add_argument("--duration", type=float, default=4.0)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--pool", type=int, default=4096)
    ap.add_argument("--ordering", choices=("plain", "learned"), default="plain")
    ap.add_argument("--model", default=None)
    args = ap.parse_args(argv)
    sat, unsat, min_rate = request_pool(args.pool, 5)
    server = SearchServer(slots=args.slots, learned=args.ordering == "learned")
    if args.ordering == "learned":
        import os
        from . import train_ordering
        train_ordering.load(args.model or os.path.join(os.path.dirname(__file__), "data", "ordering_mlp_d11.npz"))
    k = min(len(sat), len(unsat))
    pool = sat[:k] + unsat[:k]   # half satisfiable within depth 11
    print(f"pool: {len(sat)} satisfiable, {len(unsat)} not; using {2 * k}", file=sys.stderr, flush=True)
    for rate in [float(x) for x in args.rates.split(",")]:
        out = poisson_run(server, pool, rate, args.duration, min_rate, 11, seed=int(rate))
        out.update({"ordering": args.ordering, "slots": args.slots})
        print(json.dumps(out), flush=True)
    server.close()


if __name__ == "__main__":
    main()
