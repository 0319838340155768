"""Benchmark of the MI355X proof-search path (one GPU).

A step is one plain search over the whole depth-11 batch: 4,096
configurations, n uniform in 9..36, 3..12 odd cycles of lengths 3/5/7 (30 %
with an oriented 3-cycle), max_ratio 1.375, max_depth 11, dedup table of 2^27
entries, run by the library as consecutive persistent launches of at most 1,024
configurations.  The configurations are uploaded once before the timed region
and stay resident in HBM (TDP_FLAG_RESIDENT).

Prints one JSON line: states expanded per second (value), cases per hour,
the search kernel's roofline figure (algorithmic queue + dedup bytes of the
launch, from the kernel's own counters, over its HIP-event duration on the
slot's stream), the measured HBM bytes per launch (the committed rocprofv3 counter
summary's bytes per expansion times this run's expansions per launch) when
that summary is present under profiles/, the kernel's vector-ALU
issue rate against the chip's (VALU instructions per expansion from the
committed SQ-counter summary times the live expansion rate), the learned
ordering timed on the same resident batch (``learned``: states/s, cases/h
and expansions to verdict on the satisfiable configurations beside the plain
search's), and the CPU baseline (the C restatement of the reference search,
on a bounded sample of the same workload, over the host's cores, whose
decided verdicts must equal the GPU's).

``--copy-calibration MB`` instead runs a plain device copy of MB MiB, and
``--gather-calibration MB`` a gather of MB MiB of random 64-byte rows (the
known-size kernels that scale the read and write request counters).
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
# Vector-ALU issue peak of MI355X in wave64 instructions per second: 256 CUs x
# 4 SIMDs, one wave64 VALU instruction per SIMD every 4 cycles (16 lanes per
# cycle), at the 2.4 GHz peak engine clock.
VALU_PEAK_INSTS = 256 * 4 * 2.4e9 / 4
METRIC = "states expanded/s + cases/hour at depth 11, 1 MI355X; learned vs plain ordering"


def algorithmic_bytes(stats: dict, counters: dict) -> int:
    """Queue and dedup bytes of one launch, per expansion
    B = S_node (pop) + a S_node (child writes) + (a + h) B_probe + c S_node
    (a admitted, h dominated, c nodes moved between waves, each moved node
    written and read once), where B_probe = 8 B of tag per child looked up
    plus the stored key bytes compared on a tag match and the key + tag bytes
    written on an insert.  Summed over the launch from the kernel's counters."""
    nb = int(stats["node_bytes"])
    kb = int(stats["key_bytes"])
    a = int(counters["admitted"].sum())
    h = int(counters["dominated"].sum())
    moved = int(stats["donations"]) + int(stats["spills"])
    node = nb * (int(stats["pops"]) + int(stats["pushes"]) + 2 * moved)
    dedup = 8 * (a + h) + kb * int(stats["key_reads"]) + (kb + 8) * int(stats["key_writes"])
    return node + dedup


def measured_traffic(workload: str):
    """HBM bytes (read + written) per expansion of search_kernel from the
    committed rocprofv3 counter summary of this workload (the L2's memory-side
    requests counted per request size), or None when it is absent."""
    path = os.path.join(ROOT, "profiles", f"search_kernel_{workload}_hbm_bytes.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("per_expansion", {}).get("bytes")


def sq_profile(workload: str):
    """Per-expansion instruction counts of search_kernel from the committed
    rocprofv3 SQ-counter summary of this workload, or None when absent."""
    path = os.path.join(ROOT, "profiles", f"search_kernel_{workload}_sq_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def copy_calibration(mb: int) -> None:
    import torch
    n = (mb << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "dispatches": 3}))


def gather_calibration(mb: int) -> None:
    """Random 64-byte rows (16 float32, 64-byte aligned) gathered from an
    8 GiB table, MB MiB of rows per dispatch: the access pattern of the dedup
    table's tag and key lines, for scaling the read-request counters."""
    import torch
    rows_table = (8 << 30) // 64
    rows = (mb << 20) // 64
    table = torch.ones(rows_table, 16, dtype=torch.float32, device="cuda:0")
    g = torch.Generator(device="cuda:0")
    g.manual_seed(7)
    idx = torch.randint(0, rows_table, (rows,), device="cuda:0", generator=g)
    out = torch.empty(rows, 16, dtype=torch.float32, device="cuda:0")
    for _ in range(3):
        torch.index_select(table, 0, idx, out=out)
    torch.cuda.synchronize()
    print(json.dumps({"gather_rows": rows, "row_bytes": 64, "index_bytes": rows * 8,
                      "bytes_written": rows * 64, "dispatches": 3}))


def cpu_baseline(wl, sample_cases: int, cap: int, threads: int, gpu_found):
    """The oracle's search (exact dedup, one queue, children evaluated by arc
    re-linking) over the first sample_cases configurations, one configuration
    per host thread.  Every verdict it reaches within the cap must equal the
    GPU's."""
    import oracle
    cases = wl.cases[:sample_cases]
    res, secs = oracle.search_many([c.pi for c in cases], [c.spi for c in cases], [c.n for c in cases],
                                   [c.even_cycles for c in cases], [wl.min_rate] * len(cases),
                                   [wl.max_depth] * len(cases), [0] * len(cases), True, cap, threads, arcs=True)
    differ = [c for c, r in enumerate(res) if not r.aborted and r.found != bool(gpu_found[c])]
    if differ:
        raise RuntimeError(f"CPU baseline verdicts differ from the GPU's on configurations {differ[:10]}")
    exps = sum(r.counters["expansions"] for r in res)
    decided = sum(not r.aborted for r in res)
    return {"value": exps / secs, "unit": "states/s", "cores": threads, "kind": "port",
            "sample": f"first {len(cases)} configurations of the same batch, each capped at {cap} expansions "
                      f"({exps} expansions in {secs:.1f} s, {len(cases) - decided} capped); the {decided} "
                      f"verdicts reached equal the GPU's",
            "cases_per_hour_decided": decided / secs * 3600.0}


def learned_leg(T, wl, step, verdicts, args, plain_exps_sat):
    """The learned ordering (TDP_FLAG_LEARNED) over the same resident batch:
    one untimed step, then args.learned_steps timed steps; verdicts must equal
    the plain search's."""
    import numpy as np
    from tdp1375_amd import autotune, train_ordering
    train_ordering.load(args.model)
    # the learned ordering's own tuning when the cache holds one
    tuning = (autotune.apply(args.tuning, autotune.cache_key(wl.name, "learned")) if args.tuning else None)
    flags = T._lib.FLAG_LEARNED | (T._lib.FLAG_GRAPH if args.graph else 0)
    r = step(resident=True, extra=flags)
    if not np.array_equal(r.found, verdicts):
        raise RuntimeError("learned ordering changed a verdict")
    exps, kernel_ms, sat = 0, [], []
    t0 = time.perf_counter()
    for s in range(args.learned_steps):
        r = step(resident=True, extra=flags)
        if not np.array_equal(r.found, verdicts):
            raise RuntimeError("learned ordering changed a verdict")
        exps += int(r.counters["expansions"].sum())
        kernel_ms.append(T.last_launch_stats(0)["kernel_ms"])
        sat.append([int(x) for x in r.counters["expansions"][verdicts]])
        print(f"learned step {s}: {kernel_ms[-1]:.1f} ms, {int(r.counters['expansions'].sum())} expansions",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    lsat = np.array(sat[-1], np.float64)
    psat = np.array(plain_exps_sat, np.float64)
    return {"value": exps / wall, "unit": "states/s", "steps": args.learned_steps,
            "ms_per_step": 1e3 * wall / args.learned_steps,
            "cases_per_hour": len(wl.cases) * args.learned_steps / wall * 3600.0,
            "verdicts_equal_plain": True,
            "satisfiable": int(verdicts.sum()),
            "expansions_to_verdict_satisfiable": {"learned": float(lsat.sum()), "plain": float(psat.sum()),
                                                  "plain_over_learned": float(psat.sum() / max(lsat.sum(), 1.0)),
                                                  "median_learned": float(np.median(lsat)),
                                                  "median_plain": float(np.median(psat))},
            "model": os.path.relpath(args.model, ROOT),
            "tuning": tuning}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cases", type=int, default=4096)
    ap.add_argument("--queue-mb", type=int, default=32768)
    ap.add_argument("--dedup-log2", type=int, default=27)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--cpu-cap", type=int, default=100000)
    ap.add_argument("--learned-steps", type=int, default=2,
                    help="timed steps of the learned ordering on the same batch (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB")
    ap.add_argument("--gather-calibration", type=int, default=0, metavar="MB")
    ap.add_argument("--graph", action="store_true",
                    help="run each launch round as the slot's HIP graph (TDP_FLAG_GRAPH)")
    ap.add_argument("--ordering", choices=("plain", "learned"), default="plain",
                    help="learned: order children by the trained MLP in --model")
    ap.add_argument("--model", default=os.path.join(ROOT, "tdp1375_amd", "data", "ordering_mlp_d11.npz"))
    ap.add_argument("--tuning", default=os.path.join(ROOT, "tdp1375_amd", "data", "tuning.json"),
                    help="autotuner cache (python -m tdp1375_amd.autotune); its entry for this workload, "
                         "if any, is applied; '' runs the library defaults")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs every search on a single MI355X "
              "(configurations would shard across GPUs with no communication); use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gather_calibration:
        gather_calibration(args.gather_calibration)
        return

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
    from tdp1375_amd import autotune
    tuning = (autotune.apply(args.tuning, autotune.cache_key(wl.name, args.ordering)) if args.tuning else None) \
        or dict(autotune.DEFAULT)
    pis = [c.pi for c in cases]
    spis = [c.spi for c in cases]
    e0s = [c.even_cycles for c in cases]

    order_flag = 0
    if args.ordering == "learned":
        from tdp1375_amd import train_ordering
        train_ordering.load(args.model)
        order_flag = T._lib.FLAG_LEARNED
    if args.graph:
        order_flag |= T._lib.FLAG_GRAPH

    def step(resident: bool, extra=None):
        return T.search_batch(0, pis, spis, e0s, wl.min_rate, wl.max_depth, 0,
                              flags=(T._lib.FLAG_RESIDENT if resident else 0) |
                              (order_flag if extra is None else extra))

    first = None
    for w in range(max(1, args.warmup)):
        first = step(resident=w > 0)
        print(f"warmup {w}: found {int(first.found.sum())}/{len(cases)} "
              f"{T.last_launch_stats(0)['kernel_ms']:.1f} ms, {int(first.counters['expansions'].sum())} expansions",
              file=sys.stderr, flush=True)
    verdicts = first.found.copy()

    # every search_batch call returns only after its stream has drained, so
    # the timed region starts and ends with the device idle (torch is not
    # loaded here: its bundled HIP runtime would be a second one in-process)
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
        traffic.append(algorithmic_bytes(st, r.counters))
        print(f"step {s}: {st['kernel_ms']:.1f} ms, {int(r.counters['expansions'].sum())} expansions",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    plain_exps_sat = [int(x) for x in r.counters["expansions"][verdicts]]

    learned = None
    if args.learned_steps > 0 and args.ordering == "plain":
        learned = learned_leg(T, wl, step, verdicts, args, plain_exps_sat)
    T.destroy_slots()

    ms_per_step = 1e3 * wall / args.steps
    # a step runs the batch as consecutive launches of at most max_active
    # configurations; the roofline is quoted per launch (dispatch), the unit
    # of the kernel trace's average duration
    max_active = int(tuning.get("max_active") or autotune.DEFAULT["max_active"])
    dispatches = -(-len(cases) // max_active)
    avg_kernel_s = 1e-3 * sum(kernel_ms) / len(kernel_ms) / dispatches
    bytes_per_dispatch = sum(traffic) / len(traffic) / dispatches
    achieved = bytes_per_dispatch / avg_kernel_s / 1e9
    exps_per_dispatch = exps / args.steps / dispatches
    measured_per_exp = measured_traffic(wl.name) if args.ordering == "plain" else None
    cases_per_hour = len(cases) * args.steps / wall * 3600.0
    # the graph's events bracket the whole round (reset kernel, search kernel,
    # control copy), so in graph mode the time is a round's, not the kernel's
    kernel_label = "search_graph_round" if args.graph else \
        ("search_learned_kernel" if args.ordering == "learned" else "search_kernel")
    sq = sq_profile(wl.name) if args.ordering == "plain" else None
    valu = None
    if sq:
        kernel_s = 1e-3 * sum(kernel_ms)
        valu_rate = sq["valu_insts_per_expansion"] * exps / kernel_s
        valu = {"valu_insts_per_expansion": sq["valu_insts_per_expansion"],
                "achieved_insts_per_s": valu_rate, "peak_insts_per_s": VALU_PEAK_INSTS,
                "valu_frac": valu_rate / VALU_PEAK_INSTS}
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
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": wl.name, "cases": len(cases), "n_range": [9, 36], "max_ratio": wl.max_ratio,
                   "max_depth": wl.max_depth, "dedup_table_size": 1 << args.dedup_log2,
                   "queue_mb": args.queue_mb, "ordering": args.ordering, "graph": args.graph, "found": int(verdicts.sum()),
                   "tuning": tuning},
        "cases_per_hour": cases_per_hour,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_per_exp * exps_per_dispatch if measured_per_exp else None,
                     "traffic_bytes_per_expansion": measured_per_exp,
                     "algorithmic_bytes_per_expansion": bytes_per_dispatch / exps_per_dispatch,
                     "kernel": kernel_label, "dispatches_per_step": dispatches,
                     "avg_kernel_ms": 1e3 * avg_kernel_s,
                     "algorithmic_bytes_per_launch": bytes_per_dispatch,
                     "valu_frac": valu["valu_frac"] if valu else None, "valu": valu},
    }
    if learned is not None:
        out["learned"] = learned
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_sample, args.cpu_cap, threads, verdicts)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
