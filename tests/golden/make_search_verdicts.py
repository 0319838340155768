"""Writes the oracle's results for the search workloads the GPU tests check
in full (run on the CPU; the GPU tests only read the fixtures):

* ``verdicts_odd-cycle-mix-n9-36-d11.json``: the 4,096-configuration depth-11
  batch (n uniform in 9..36, max_ratio 1.375, D = 11).  One character per
  configuration: ``1`` found, ``0`` not found, ``?`` undecided because the
  oracle reached the per-configuration expansion budget first.
* ``search_four-3cycles-n24.npz`` and ``search_four-3cycles-n24-unoriented.npz``:
  the 2,048 n = 24 configurations (four 3-cycles, D = 9), each with the
  oracle's verdict, sequence (positions) and counters of its deterministic
  search, which the GPU's deterministic mode reproduces exactly.

The oracle is the C restatement in oracle/ (exact dedup, one queue), run with
arc re-linking for the children (``arcs=True``), whose results and counters
equal the literal recount's (tests/test_oracle.py).

    python tests/golden/make_search_verdicts.py [--threads 7] [--budget 2000000] [--only c3|c2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tdp1375_amd import workloads as W  # noqa: E402

COUNTERS = ("expansions", "generated", "admitted", "dominated", "improved", "inserted", "roots_live")


def run(wl, budget: int, threads: int, chunk: int = 256):
    cases = wl.cases
    out = []
    t0 = time.time()
    for base in range(0, len(cases), chunk):
        part = cases[base:base + chunk]
        res, _ = oracle.search_many([c.pi for c in part], [c.spi for c in part], [c.n for c in part],
                                    [c.even_cycles for c in part], [wl.min_rate] * len(part),
                                    [wl.max_depth] * len(part), [0] * len(part), True, budget, threads,
                                    arcs=True)
        out.extend(res)
        print(f"{wl.name}: {base + len(part)}/{len(cases)} in {time.time() - t0:.0f} s, "
              f"{sum(r.aborted for r in out)} past the budget", flush=True)
    return out


def write_verdicts(wl, res, budget: int, path: str):
    verdicts = "".join("?" if r.aborted else ("1" if r.found else "0") for r in res)
    doc = {"workload": wl.name, "cases": len(res), "max_ratio": wl.max_ratio, "max_depth": wl.max_depth,
           "min_rate_f32": wl.min_rate, "budget_expansions": budget,
           "generator": "tdp1375_amd.workloads.workload_odd_cycle_mix_d11",
           "found": verdicts.count("1"), "not_found": verdicts.count("0"), "undecided": verdicts.count("?"),
           "verdicts": verdicts}
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=0)
        fh.write("\n")


def write_full(wl, res, path: str):
    found = np.array([r.found for r in res], np.uint8)
    depth = np.array([r.depth for r in res], np.int8)
    moves = np.full((len(res), oracle.MAX_DEPTH, 3), -1, np.int8)
    for c, r in enumerate(res):
        for m, mv in enumerate(r.moves):
            moves[c, m] = mv
    counters = np.array([[r.counters[k] for k in COUNTERS] for r in res], np.uint64)
    if any(r.aborted for r in res):
        raise RuntimeError("a configuration reached the budget: the fixture needs complete searches")
    np.savez_compressed(path, found=found, depth=depth, moves=moves, counters=counters,
                        counter_names=np.array(COUNTERS), max_depth=np.int32(wl.max_depth),
                        min_rate=np.float32(wl.min_rate))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=max(1, (os.cpu_count() or 2) - 1))
    ap.add_argument("--budget", type=int, default=2_000_000,
                    help="expansions per configuration of the depth-11 batch before it counts as undecided")
    ap.add_argument("--only", choices=("c2", "c3"), default=None)
    a = ap.parse_args()
    if a.only in (None, "c2"):
        for oriented in (True, False):
            wl = W.workload_four_3cycles_n24(oriented)
            res = run(wl, 0, a.threads)
            write_full(wl, res, os.path.join(HERE, f"search_{wl.name}.npz"))
    if a.only in (None, "c3"):
        wl = W.workload_odd_cycle_mix_d11(4096)
        res = run(wl, a.budget, a.threads, chunk=128)
        write_verdicts(wl, res, a.budget, os.path.join(HERE, f"verdicts_{wl.name}.json"))


if __name__ == "__main__":
    main()
