"""Train the move-ordering MLP on the GPU from the plain search's own sequences.

    python -m tdp1375_amd.train_ordering --out tdp1375_amd/data/ordering_mlp_d11.npz

Training pairs: the sequences the plain search returns for the two variants of
the n=24 four-3-cycle batch (2,048 configurations each, max_depth 9) and for
held-out batches of the depth-11 generator (4,096 configurations each, seeds
11..18 by default; the benchmark batch uses seed 3).  Every state on a
returned sequence is a group of the children the search admits there,
labelled by the move the sequence took (learned.build_dataset, built on
--workers host processes); tdp_mlp_train fits the MLP on the GPU.

Networks trained on the same data differ (initialisation and batch order,
--seed, --seed + 1, ...), and so do the learned search's times with them by
10-40 %.  --candidates networks are trained and each is run as the learned
ordering over a validation batch of the depth-11 generator (seed 19 by
default, in neither the training nor the benchmark batches, with the cached
learned-ordering tuning); the one with the least median search time is kept.
Its weights and a JSON summary of the run (every candidate's validation
figures included) are written to --out.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

from . import autotune
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


def validate(cases, min_rate, max_depth, verdicts, steps: int):
    """Median kernel time (ms) and expansions of the learned search over
    `cases` with the current network; its verdicts must equal `verdicts`
    (the plain search's)."""
    pis, spis, e0s = [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases]
    times, exps = [], []
    for step in range(steps):
        flags = G._lib.FLAG_LEARNED | (G._lib.FLAG_RESIDENT if step else 0)
        res = G.search_batch(0, pis, spis, e0s, min_rate, max_depth, 0, flags=flags)
        if not np.array_equal(res.found, verdicts):
            raise RuntimeError("learned verdicts differ from the plain search's on the validation batch")
        times.append(G.last_launch_stats(0)["kernel_ms"])
        exps.append(int(res.counters["expansions"].sum()))
    return float(np.median(times)), float(np.median(exps))


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
    ap.add_argument("--seed", type=int, default=8)
    ap.add_argument("--seeds", default="11,12,13,14,15,16,17,18",
                    help="generator seeds of the depth-11 batches (the benchmark uses 3)")
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--candidates", type=int, default=5, help="networks trained (training seeds --seed, --seed + 1, ...)")
    ap.add_argument("--validation-seed", type=int, default=19,
                    help="generator seed of the depth-11 batch that picks among the candidates")
    ap.add_argument("--validation-steps", type=int, default=5)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[{time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)

    seeds = [int(x) for x in args.seeds.split(",") if x]
    if 3 in seeds or args.validation_seed == 3:
        raise SystemExit("seed 3 is the benchmark batch; train and validate on other seeds")
    if args.validation_seed in seeds:
        raise SystemExit("the validation seed must not be a training seed")
    if args.candidates < 1:
        raise SystemExit("--candidates must be at least 1")
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
    d11 = W.workload_odd_cycle_mix_d11(1)
    val = W.odd_cycle_mix(4096, args.validation_seed)
    val_plain, _ = plain_sequences(val, d11.min_rate, d11.max_depth, 0)
    autotune.apply(autotune.DEFAULT_CACHE, autotune.cache_key(d11.name, "learned"))
    best, candidates = None, []
    for k in range(args.candidates):
        seed = args.seed + k
        LN.init(seed)
        t0 = time.time()
        losses = LN.train(data, epochs=args.epochs, lr=args.lr, groups_per_batch=args.batch_groups, seed=seed)
        train_s = time.time() - t0
        ms, exps = validate(val, d11.min_rate, d11.max_depth, val_plain.found, args.validation_steps)
        cand = {"seed": seed, "train_seconds": train_s, "epoch_loss": [float(x) for x in losses],
                "group_accuracy": LN.group_accuracy(data, LN.forward(data.rows)),
                "validation_kernel_ms": ms, "validation_expansions": exps}
        candidates.append(cand)
        log(f"candidate {k} (seed {seed}): loss {losses[0]:.3f} -> {losses[-1]:.3f}, accuracy "
            f"{cand['group_accuracy']:.3f}, validation {ms:.1f} ms, {exps:.3g} expansions")
        if best is None or ms < best[0]["validation_kernel_ms"]:
            best = (cand, LN.get_weights())
    cand, (w1, b1, w2, b2) = best
    summary.update({"groups": data.groups, "rows": int(len(data.rows)), "epochs": args.epochs, "lr": args.lr,
                    "groups_per_batch": args.batch_groups, "seed": cand["seed"],
                    "train_seconds": cand["train_seconds"], "epoch_loss": cand["epoch_loss"],
                    "group_accuracy": cand["group_accuracy"],
                    "validation": {"workload": f"odd-cycle-mix-n9-36-d11 seed {args.validation_seed}",
                                   "steps": args.validation_steps, "found": int(val_plain.found.sum())},
                    "candidates": candidates})
    log(f"kept the network of seed {cand['seed']}")
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
