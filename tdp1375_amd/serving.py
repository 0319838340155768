"""Serving: single-configuration search requests through the device-resident
request loop (tdp_serve_*).

The reference serves the proof generator's ForkJoinPool: every worker thread
takes a slot, runs one blocking search and gives the slot back, one launch per
call (GPUSortingSearch.java:19-28,47-87, SortOrExtend.java:96-104).  Here one
persistent launch serves every caller: its waves take requests as they arrive
(continuous batching, no launch boundaries), hand work to idle waves, share
the queue memory and the dedup table, and reply through pinned host memory;
a request waits for admission only while no device slot is free or its root
nodes would not fit in the queue memory beside the requests in flight.

DeviceServer wraps a session.  Its submit() hands a request to a pool of
caller threads (the ForkJoin workers' stand-in), each of which blocks in
tdp_serve_search until its own reply, and returns a Future.  A request's
verdict does not depend on what it runs beside: that only changes which wave
takes which node first.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from . import _lib
from . import workloads as W


@dataclass
class Request:
    pi: Sequence[int]
    spi: Sequence[int]
    even_cycles: int
    min_rate: float
    max_depth: int
    arrival: float = 0.0


@dataclass
class Reply:
    found: bool
    moves: list            # positions (i, j, k) per move
    expansions: int        # of the batches ended before the verdict
    latency: float         # seconds from arrival to the verdict


class DeviceServer:
    """A device-resident serving session on the current GPU."""

    def __init__(self, queue_bytes: int = 16 << 30, dedup: int = 1 << 26, slots: int = 256, max_n: int = 63,
                 learned: bool = False, lifetime_s: int = 600, callers: int = 0):
        self.lib = _lib.load()
        flags = _lib.FLAG_LEARNED if learned else 0
        _lib.check(self.lib.tdp_serve_start(queue_bytes, dedup, slots, max_n, flags, lifetime_s))
        # one caller thread per request the session admits at once
        self._pool = ThreadPoolExecutor(max_workers=callers or slots, thread_name_prefix="serve-caller")
        self._closed = False
        self._lock = threading.Lock()

    def search(self, pi, spi, e0: int, min_rate: float, max_depth: int, arrival: float = 0.0) -> Reply:
        """One blocking request (the caller thread waits in the library)."""
        n = len(pi)
        moves = (ctypes.c_int32 * (_lib.MAX_DEPTH * 3))()
        depth = ctypes.c_int32(0)
        cnt = _lib.Counters()
        t0 = arrival or time.perf_counter()
        st = self.lib.tdp_serve_search(_lib.u8(pi, 64), _lib.u8(spi, 64), n, e0, ctypes.c_float(min_rate),
                                       max_depth, 0, moves, ctypes.byref(depth), ctypes.byref(cnt))
        done = time.perf_counter()
        _lib.check(st)
        d = depth.value
        pos = [(moves[3 * m], moves[3 * m + 1], moves[3 * m + 2]) for m in range(d)]
        return Reply(d > 0, pos, int(cnt.expansions), done - t0)

    def submit(self, req: Request) -> Future:
        if not req.arrival:
            req.arrival = time.perf_counter()
        return self._pool.submit(self.search, req.pi, req.spi, req.even_cycles, req.min_rate, req.max_depth,
                                 req.arrival)

    def stats(self) -> dict:
        c, f, r = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int64()
        _lib.check(self.lib.tdp_serve_stats(ctypes.byref(c), ctypes.byref(f), ctypes.byref(r)))
        return {"completed": c.value, "in_flight": f.value, "reserved_bytes": r.value}

    def close(self) -> None:
        with self._lock:
            if self._closed:
                return
            self._closed = True
        self._pool.shutdown(wait=True)
        _lib.check(self.lib.tdp_serve_stop())

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def poisson_run(server: DeviceServer, pool: List, rate_per_s: float, duration_s: float, min_rate: float,
                max_depth: int, seed: int) -> dict:
    """Poisson arrivals for duration_s; returns throughput and latency
    percentiles (arrival to verdict) and the replies in arrival order with
    the pool index of each request."""
    rng = np.random.default_rng(seed)
    futures, picks = [], []
    t0 = time.perf_counter()
    t_next = t0
    while True:
        t_next += rng.exponential(1.0 / rate_per_s)
        if t_next - t0 > duration_s:
            break
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        k = int(rng.integers(len(pool)))
        c = pool[k]
        picks.append(k)
        futures.append(server.submit(Request(c.pi, c.spi, c.even_cycles, min_rate, max_depth, arrival=t_next)))
    replies = [f.result() for f in futures]
    wall = time.perf_counter() - t0
    lat = np.array([r.latency for r in replies]) * 1e3
    return {"offered_rate": rate_per_s, "requests": len(replies), "throughput": len(replies) / wall,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "found": int(sum(r.found for r in replies)),
            "_replies": replies, "_picks": picks}


def pool_cases(count: int = 1024, seed: int = 5) -> List[W.Case]:
    """The request generator: n uniform in 9..48, the depth-11 generator's
    odd cycles (3..12 cycles of lengths 3/5/7, 30 % with an oriented 3-cycle)."""
    return W.odd_cycle_mix(count, seed, 9, 48)


def main(argv=None):
    import os
    ap = argparse.ArgumentParser(description="Poisson serving sweep through the device-resident loop")
    ap.add_argument("--rates", default="10,100,1000,10000")
    ap.add_argument("--duration", type=float, default=4.0)
    ap.add_argument("--ordering", choices=("plain", "learned"), default="plain")
    ap.add_argument("--model", default=None)
    ap.add_argument("--share-interval", type=int, default=None,
                    help="expansions between hand-off checks (default: 2 for the learned ordering, whose "
                         "requests settle in a few expansions and leave hard ones to spread, else the "
                         "library default)")
    ap.add_argument("--pool", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                   "tests", "golden", "serving_pool_n9-48-d11.json"),
                    help="labelled request pool (indices into pool_cases and their verdicts)")
    args = ap.parse_args(argv)
    with open(args.pool) as fh:
        doc = json.load(fh)
    cases = pool_cases(doc["generated"], doc["seed"])
    pool = [cases[i] for i in doc["satisfiable"]] + [cases[i] for i in doc["unsatisfiable"]]
    rate = float(np.float32(2.0) / np.float32(doc["max_ratio"]))
    if args.ordering == "learned":
        from . import train_ordering
        train_ordering.load(args.model or os.path.join(os.path.dirname(__file__), "data", "ordering_mlp_d11.npz"))
    share_interval = args.share_interval if args.share_interval is not None else (2 if args.ordering == "learned" else 0)
    if share_interval:
        from . import gpu_sorting_search
        gpu_sorting_search.set_tuning(share_interval=share_interval)
    with DeviceServer(learned=args.ordering == "learned", max_n=48) as server:
        for r in [float(x) for x in args.rates.split(",")]:
            out = poisson_run(server, pool, r, args.duration, rate, doc["max_depth"], seed=int(r))
            out = {k: v for k, v in out.items() if not k.startswith("_")}
            out.update({"ordering": args.ordering})
            print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
