"""Benchmark of the MI355X proof-search path (one GPU).

A step is one persistent launch of the plain search over the whole depth-11
batch: 4,096 configurations, n uniform in 9..36, 3..12 odd cycles of lengths
3/5/7 (30 % with an oriented 3-cycle), max_ratio 1.375, max_depth 11, dedup
table of 2^27 entries.  The configurations are uploaded once before the timed
region and stay resident in HBM (TDP_FLAG_RESIDENT).

Prints one JSON line: states expanded per second (value), cases per hour,
the search kernel's HBM roofline figure (algorithmic bytes from the kernel's
own traffic counters over its HIP-event duration) and the CPU baseline (the
C restatement of the reference search, on a bounded sample of the same
workload, over the host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "states expanded/s + cases/hour at depth 11, 1 MI355X; learned vs plain ordering"


def algorithmic_bytes(stats: dict) -> int:
    """Queue and dedup traffic one launch has to move:
    node reads (pops + steals) and writes (pushes + donations, each donation
    also reads the node); 8-byte dedup tags read (probes), stored keys
    compared (key bytes plus the 8-byte tag re-read), and entries written
    (key bytes, the tag CAS and the tag publication, 8 + 8 bytes)."""
    nb = int(stats["node_bytes"])
    kb = int(stats["key_bytes"])
    node = nb * (int(stats["pops"]) + int(stats["pushes"]) + 2 * int(stats["donations"]) + int(stats["steals"]))
    dedup = 8 * int(stats["probes"]) + (kb + 8) * int(stats["key_reads"]) + (kb + 16) * int(stats["key_writes"])
    return node + dedup


def cpu_baseline(wl, sample_cases: int, cap: int, threads: int):
    import oracle
    cases = wl.cases[:sample_cases]
    res, secs = oracle.search_many([c.pi for c in cases], [c.spi for c in cases], [c.n for c in cases],
                                   [c.even_cycles for c in cases], [wl.min_rate] * len(cases),
                                   [wl.max_depth] * len(cases), [0] * len(cases), True, cap, threads)
    exps = sum(r.counters["expansions"] for r in res)
    return {"value": exps / secs, "unit": "states/s", "cores": threads, "kind": "port",
            "sample": f"first {len(cases)} configurations of the same batch, each capped at {cap} expansions "
                      f"({exps} expansions in {secs:.1f} s, {sum(r.aborted for r in res)} capped)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cases", type=int, default=4096)
    ap.add_argument("--queue-mb", type=int, default=32768)
    ap.add_argument("--dedup-log2", type=int, default=27)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=48)
    ap.add_argument("--cpu-cap", type=int, default=20000)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs every search on a single MI355X "
              "(configurations would shard across GPUs with no communication); use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import tdp1375_amd as T
    from tdp1375_amd import workloads as W

    lib = T.load()
    if lib.tdp_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    wl = W.workload_odd_cycle_mix_d11(args.cases)
    cases = wl.cases
    T.init_slots(1, 1, args.queue_mb << 20, 1 << args.dedup_log2)
    pis = [c.pi for c in cases]
    spis = [c.spi for c in cases]
    e0s = [c.even_cycles for c in cases]

    def step(resident: bool):
        return T.search_batch(0, pis, spis, e0s, wl.min_rate, wl.max_depth, 0,
                              flags=T._lib.FLAG_RESIDENT if resident else 0)

    first = None
    for w in range(max(1, args.warmup)):
        first = step(resident=w > 0)
        print(f"warmup {w}: found {int(first.found.sum())}/{len(cases)} "
              f"{T.last_launch_stats(0)['kernel_ms']:.1f} ms", file=sys.stderr, flush=True)
    verdicts = first.found.copy()

    exps = 0
    kernel_ms = []
    traffic = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        r = step(resident=True)
        st = T.last_launch_stats(0)
        if not np.array_equal(r.found, verdicts):
            raise RuntimeError("verdicts changed between steps")
        exps += int(r.counters["expansions"].sum())
        kernel_ms.append(st["kernel_ms"])
        traffic.append(algorithmic_bytes(st))
        print(f"step {s}: {st['kernel_ms']:.1f} ms, {int(r.counters['expansions'].sum())} expansions",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    T.destroy_slots()

    ms_per_step = 1e3 * wall / args.steps
    avg_kernel_s = 1e-3 * sum(kernel_ms) / len(kernel_ms)
    achieved = (sum(traffic) / len(traffic)) / avg_kernel_s / 1e9
    cases_per_hour = len(cases) * args.steps / wall * 3600.0
    out = {
        "metric": METRIC,
        "value": exps / wall,
        "unit": "states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": wl.name, "cases": len(cases), "n_range": [9, 36], "max_ratio": wl.max_ratio,
                   "max_depth": wl.max_depth, "dedup_table_size": 1 << args.dedup_log2,
                   "queue_mb": args.queue_mb, "ordering": "plain", "found": int(verdicts.sum())},
        "cases_per_hour": cases_per_hour,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "search_kernel", "avg_kernel_ms": 1e3 * avg_kernel_s},
    }
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_sample, args.cpu_cap, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
