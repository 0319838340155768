"""Writes serving_pool_n9-48-d11.json: the request pool of the serving tests
and sweep, labelled by the oracle (run on the CPU; the GPU tests only read it).

Candidates are the first `generated` configurations of the serving generator
(tdp1375_amd.serving.pool_cases: n uniform in 9..48, the depth-11 generator's
odd cycles, seed 5).  The oracle's search (exact dedup, one queue, children
by arc re-linking) runs each at max_ratio 1.375 and D = 11 with an expansion
budget; the pool is the first `per_class` satisfiable and the first
`per_class` unsatisfiable configurations it decides within the budget, so
that half of the pool is satisfiable and every label is the oracle's.

    python tests/golden/make_serving_pool.py [--threads 8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tdp1375_amd import serving as S  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=max(1, os.cpu_count() or 1))
    ap.add_argument("--generated", type=int, default=4096)
    ap.add_argument("--per-class", type=int, default=128)
    ap.add_argument("--budget", type=int, default=60000)
    a = ap.parse_args()
    seed, max_ratio, max_depth = 5, 1.375, 11
    cases = S.pool_cases(a.generated, seed)
    rate = float(np.float32(2.0) / np.float32(max_ratio))
    res, secs = oracle.search_many([c.pi for c in cases], [c.spi for c in cases], [c.n for c in cases],
                                   [c.even_cycles for c in cases], [rate] * len(cases), [max_depth] * len(cases),
                                   [0] * len(cases), True, a.budget, a.threads, arcs=True)
    sat = [i for i, r in enumerate(res) if not r.aborted and r.found][:a.per_class]
    unsat = [i for i, r in enumerate(res) if not r.aborted and not r.found][:a.per_class]
    doc = {"generator": "tdp1375_amd.serving.pool_cases", "seed": seed, "generated": a.generated,
           "max_ratio": max_ratio, "max_depth": max_depth, "budget_expansions": a.budget,
           "undecided_within_budget": sum(r.aborted for r in res), "oracle_seconds": round(secs, 1),
           "satisfiable": sat, "unsatisfiable": unsat,
           "n": {str(i): cases[i].n for i in sat + unsat}}
    with open(os.path.join(HERE, "serving_pool_n9-48-d11.json"), "w") as fh:
        json.dump(doc, fh)
        fh.write("\n")
    print(f"{len(sat)} satisfiable, {len(unsat)} unsatisfiable, "
          f"{doc['undecided_within_budget']} undecided, {secs:.0f} s")


if __name__ == "__main__":
    main()
