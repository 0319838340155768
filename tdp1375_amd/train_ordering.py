"""Train the move-ordering MLP on the GPU from the plain search's own sequences.

    python -m tdp1375_amd.train_ordering --out tdp1375_amd/data/ordering_mlp_d11.npz

Training pairs: the sequences the plain search returns for the two variants of
the n=24 four-3-cycle batch (2,048 configurations each, max_depth 9) and for
held-out batches of the depth-11 generator (4,096 configurations each, seeds
11..18 by default; the benchmark batch uses seed 3).  Every state on a
returned sequence is a group of the children the search admits there,
labelled by the move the sequence took (learned.build_dataset, built on
--workers host processes); tdp_mlp_train fits the MLP on the GPU.  The
weights and a JSON summary of the run are written to --out.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

from . import gpu_sorting_search as G
from . import learned as LN
from . import workloads as W


def plain_sequences(cases, min_rate, max_depth, flags):
    res = G.search_batch(0, [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases], min_rate,
                         max_depth, 0, flags=flags)
    return res, [res.moves[c] if res.found[c] else [] for c in range(len(cases))]


def _build(args):
    cases, min_rate, max_depth, seqs = args
    return LN.build_dataset(cases, min_rate, max_depth, seqs)


def build_parallel(pool, cases, min_rate, max_depth, seqs, parts: int = 64):
    """learned.build_dataset over `parts` slices of the cases on the pool's
    processes, concatenated in case order."""
    idx = np.array_split(np.arange(len(cases)), parts)
    dss = pool.map(_build, [([cases[i] for i in ix], min_rate, max_depth, [seqs[i] for i in ix]) for ix in idx])
    return concat([d for d in dss if d.groups > 0])


def concat(sets):
    starts, off = [np.zeros(1, np.int32)], 0
    for d in sets:
        starts.append(d.group_start[1:] + off)
        off += int(d.group_start[-1])
    return LN.Dataset(np.concatenate([d.rows for d in sets]), np.concatenate([d.labels for d in sets]),
                      np.concatenate(starts).astype(np.int32))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--batch-groups", type=int, default=64)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--seeds", default="11,12,13,14,15,16,17,18",
                    help="generator seeds of the depth-11 batches (the benchmark uses 3)")
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[{time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)

    seeds = [int(x) for x in args.seeds.split(",") if x]
    if 3 in seeds:
        raise SystemExit("seed 3 is the benchmark batch; train on other seeds")
    G.init_slots(1, 1, 32768 << 20, 1 << 27)
    sets, summary = [], {"sources": []}
    with Pool(max(1, args.workers)) as pool:
        for oriented in (False, True):
            wl = W.workload_four_3cycles_n24(oriented)
            res, seqs = plain_sequences(wl.cases, wl.min_rate, wl.max_depth, G._lib.FLAG_DETERMINISTIC)
            ds = build_parallel(pool, wl.cases, wl.min_rate, wl.max_depth, seqs)
            sets.append(ds)
            summary["sources"].append({"workload": wl.name,
                                       "found": int(res.found.sum()), "groups": ds.groups, "rows": int(len(ds.rows))})
            log(f"{summary['sources'][-1]}")
        d11 = W.workload_odd_cycle_mix_d11(1)
        for sd in seeds:
            held = W.odd_cycle_mix(4096, sd)
            res, seqs = plain_sequences(held, d11.min_rate, d11.max_depth, 0)
            ds = build_parallel(pool, held, d11.min_rate, d11.max_depth, seqs)
            sets.append(ds)
            summary["sources"].append({"workload": f"odd-cycle-mix-n9-36-d11 seed {sd} (held out)",
                                       "found": int(res.found.sum()), "groups": ds.groups, "rows": int(len(ds.rows))})
            log(f"{summary['sources'][-1]}")
    data = concat(sets)
    LN.init(args.seed)
    t0 = time.time()
    losses = LN.train(data, epochs=args.epochs, lr=args.lr, groups_per_batch=args.batch_groups, seed=args.seed)
    summary.update({"groups": data.groups, "rows": int(len(data.rows)), "epochs": args.epochs, "lr": args.lr,
                    "groups_per_batch": args.batch_groups, "train_seconds": time.time() - t0,
                    "epoch_loss": [float(x) for x in losses],
                    "group_accuracy": LN.group_accuracy(data, LN.forward(data.rows))})
    log(f"trained: loss {losses[0]:.3f} -> {losses[-1]:.3f}, accuracy {summary['group_accuracy']:.3f}")
    w1, b1, w2, b2 = LN.get_weights()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, w1=w1, b1=b1, w2=w2, b2=np.float32(b2))
    with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
        json.dump(summary, fh, indent=1)
    G.destroy_slots()


def load(path: str) -> None:
    """Sets the model of this process from a file written by main()."""
    with np.load(path, allow_pickle=False) as z:
        LN.set_weights(z["w1"], z["b1"], z["w2"], float(z["b2"]))


if __name__ == "__main__":
    main()
