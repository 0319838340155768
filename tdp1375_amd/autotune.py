"""Autotuner of the work-sharing launch.

Chooses the knobs the kernel cannot pick for itself on a given batch: waves
per CU (queue depth per wave follows from it and the queue budget), nodes
per hand-off, expansions between hand-off checks, nodes per expansion batch
and configurations per persistent launch.  It runs the batch once to upload
it, then re-runs it from device memory (TDP_FLAG_RESIDENT) once per trial,
one knob at a time from the current best (coordinate descent), and keeps the
setting with the least kernel time.  Every trial must reach the same
verdicts; a trial that does not is an error, never a candidate.

The result is cached in a JSON file keyed by the workload name, and
``apply(cache, key)`` sets a cached tuning before a run (bench.py reads
tdp1375_amd/data/tuning.json this way).

    python -m tdp1375_amd.autotune --out tdp1375_amd/data/tuning.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from . import gpu_sorting_search as G

KNOBS = ("waves_per_cu", "share_interval", "share_max", "batch", "max_active")
DEFAULT = {"waves_per_cu": 0, "share_interval": 8, "share_max": 8, "batch": 12, "max_active": 1024}
GRID = {
    "waves_per_cu": [0, 8],                # 0: the kernel's occupancy limit
    "share_interval": [8, 16, 32],
    "share_max": [4, 8, 16],
    "batch": [4, 8, 12, 16],
    "max_active": [512, 1024, 1536, 2048, 3072],
}
# The learned ordering settles most configurations within a few expansions
# and leaves a tail of hard ones, each expansion costing more (the scoring):
# more frequent and larger hand-offs and larger launches are worth trying.
GRID_LEARNED = {
    "waves_per_cu": [0, 8],
    "share_interval": [2, 4, 8, 16],
    "share_max": [8, 16, 32],
    "batch": [4, 8],
    "max_active": [1024, 2048, 4096],
}
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning.json")


def set_knobs(k: Dict[str, int]) -> None:
    G.set_tuning(waves_per_cu=k["waves_per_cu"], share_interval=k["share_interval"], share_max=k["share_max"],
                 max_active=k["max_active"])
    G.set_expansion_batch(k["batch"])


def reset() -> None:
    G.set_tuning()
    G.set_expansion_batch(0)


def autotune(slot: int, pis, spis, e0s, min_rate, max_depth, grid: Optional[Dict[str, List[int]]] = None,
             start: Optional[Dict[str, int]] = None, repeats: int = 1, flags: int = 0, log=None,
             passes: int = 1) -> dict:
    """Coordinate descent over grid from start (up to `passes` sweeps over
    the knobs, stopping after a sweep that changes nothing); returns
    {"best": knobs, "kernel_ms": t, "trials": [...]} and leaves the best
    tuning set."""
    grid = grid or GRID
    best = dict(start or DEFAULT)
    trials = []
    reference = None

    def run(k):
        nonlocal reference
        set_knobs(k)
        times = []
        for _ in range(repeats):
            res = G.search_batch(slot, pis, spis, e0s, min_rate, max_depth, 0, flags=flags | _lib.FLAG_RESIDENT)
            if reference is None:
                reference = res.found.copy()
            elif not np.array_equal(reference, res.found):
                raise RuntimeError(f"autotune: verdicts changed under tuning {k}")
            times.append(G.last_launch_stats(slot)["kernel_ms"])
        # the parallel search's expansion count varies run to run: the
        # median over the repeats, not the luckiest run
        t = float(np.median(times))
        trials.append({"knobs": dict(k), "kernel_ms": t})
        if log:
            log(f"{k} -> {t:.1f} ms")
        return t

    # upload once, then every trial runs from device memory
    set_knobs(best)
    first = G.search_batch(slot, pis, spis, e0s, min_rate, max_depth, 0, flags=flags)
    reference = first.found.copy()
    best_t = run(best)
    measured = {tuple(sorted(best.items()))}
    for _ in range(max(1, passes)):
        before = dict(best)
        for knob in KNOBS:
            for v in grid.get(knob, []):
                k = dict(best)
                k[knob] = v
                if tuple(sorted(k.items())) in measured:
                    continue
                measured.add(tuple(sorted(k.items())))
                t = run(k)
                if t < best_t:
                    best, best_t = k, t
        if best == before:
            break
    set_knobs(best)
    return {"best": best, "kernel_ms": best_t, "trials": trials}


def save(path: str, key: str, result: dict, note: str = "") -> None:
    data = {}
    if os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
    data[key] = {"best": result["best"], "kernel_ms": result["kernel_ms"], "trials": result["trials"], "note": note}
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1)
    os.replace(tmp, path)


def cache_key(workload: str, ordering: str = "plain") -> str:
    """Cache entry of a workload under an ordering (the plain one keeps the
    workload's own name)."""
    return workload if ordering == "plain" else f"{workload}/{ordering}"


def apply(path: str, key: str) -> Optional[Dict[str, int]]:
    """Sets the cached tuning of key (if any) and returns it."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        entry = json.load(fh).get(key)
    if not entry:
        return None
    set_knobs(entry["best"])
    return entry["best"]


def main(argv=None):
    from . import workloads as W
    ap = argparse.ArgumentParser(description="autotune the work-sharing launch on the depth-11 batch")
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--cases", type=int, default=4096)
    ap.add_argument("--queue-mb", type=int, default=32768)
    ap.add_argument("--dedup-log2", type=int, default=27)
    ap.add_argument("--repeats", type=int, default=1)
    ap.add_argument("--passes", type=int, default=1, help="coordinate-descent sweeps over the knobs")
    ap.add_argument("--ordering", choices=("plain", "learned"), default="plain")
    ap.add_argument("--model", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                    "ordering_mlp_d11.npz"))
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[{time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)

    wl = W.workload_odd_cycle_mix_d11(args.cases)
    flags, grid = 0, GRID
    if args.ordering == "learned":
        from . import train_ordering
        train_ordering.load(args.model)
        flags, grid = _lib.FLAG_LEARNED, GRID_LEARNED
    G.init_slots(1, 1, args.queue_mb << 20, 1 << args.dedup_log2)
    try:
        res = autotune(0, [c.pi for c in wl.cases], [c.spi for c in wl.cases], [c.even_cycles for c in wl.cases],
                       wl.min_rate, wl.max_depth, grid=grid, repeats=args.repeats, flags=flags, log=log,
                       passes=args.passes)
    finally:
        reset()
        G.destroy_slots()
    key = cache_key(wl.name, args.ordering)
    save(args.out, key, res, note=f"{args.cases} configurations, {args.ordering} ordering, queue {args.queue_mb} MiB, "
                                  f"dedup 2^{args.dedup_log2}, coordinate descent from the library defaults "
                                  f"({args.passes} pass(es) at most, {args.repeats} run(s) per trial)")
    print(json.dumps({"workload": key, "best": res["best"], "kernel_ms": res["kernel_ms"],
                      "trials": len(res["trials"])}), flush=True)


if __name__ == "__main__":
    main()
