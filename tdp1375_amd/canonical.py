"""Configuration canonicalisation on the GPU (host-side mirror).

Mirrors the part of Configuration.java that keys proof cases
(Configuration.java:53-198,243-297): a configuration is spi (a product of
cycles, no fixed points) over pi (a cycle, the canonical 0..n-1 unless
given); its signature labels every position of pi by the cycle of its symbol
(labels by first appearance, oriented cycles with a rank in hundredths); the
canonical configuration is the one of the least of the 2n signatures of
every rotation of pi and its mirror, by CRC32 hash then content, and that
hash is the configuration's hashCode.  The work runs batched in one launch
(tdp_canonical_batch).
"""
from __future__ import annotations

import ctypes
import re
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib

MAX_N = 64


def parse_cycles(text: str) -> List[List[int]]:
    """'(0 9 7)(1 5 2 6 3)' -> [[0, 9, 7], [1, 5, 2, 6, 3]] (MulticyclePermutation.java:78-88)."""
    return [[int(x) for x in re.split(r"[,\s]+", g.strip()) if x] for g in re.findall(r"\(([^()]*?)\)", text)]


def one_line(cycles: Sequence[Sequence[int]], n: int) -> List[int]:
    img = list(range(n))
    for c in cycles:
        for t, s in enumerate(c):
            img[s] = c[(t + 1) % len(c)]
    return img


def cycles_of(img: Sequence[int]) -> List[List[int]]:
    seen, out = set(), []
    for s in range(len(img)):
        if s in seen:
            continue
        c, x = [], s
        while x not in seen:
            seen.add(x)
            c.append(x)
            x = img[x]
        out.append(c)
    return out


def format_cycles(cycles: Sequence[Sequence[int]]) -> str:
    return "".join("(" + " ".join(str(s) for s in c) + ")" for c in cycles)


@dataclass
class Canonical:
    signature: np.ndarray       # float32[n]: the canonical signature
    hash: int                   # Configuration.hashCode
    index: int                  # 2 * rotation + mirror of the equivalent signature it is
    spi: List[int]              # canonical configuration, one-line (pi the identity)

    @property
    def cycles(self) -> List[List[int]]:
        return cycles_of(self.spi)


def canonical_batch(spis: Sequence[Sequence[int]], pis: Optional[Sequence[Sequence[int]]] = None) -> List[Canonical]:
    """Canonical forms of many configurations in one launch.  spis[c] is a
    one-line spi (image of every symbol); pis[c] the symbols of pi by position
    (default 0..n-1)."""
    count = len(spis)
    if count == 0:
        return []
    S = np.zeros((count, MAX_N), np.uint8)
    P = np.zeros((count, MAX_N), np.uint8)
    ns = np.zeros(count, np.int32)
    for c, spi in enumerate(spis):
        n = len(spi)
        if n > MAX_N:
            raise _lib.TdpError(_lib.TDP_ERR_INVALID_ARGUMENT, "configuration size must be in 1..64")
        ns[c] = n
        S[c, :n] = np.asarray(spi, np.uint8)
        P[c, :n] = np.asarray(pis[c] if pis is not None else range(n), np.uint8)
    sig = np.zeros((count, MAX_N), np.float32)
    h = np.zeros(count, np.int32)
    idx = np.zeros(count, np.int32)
    out = np.zeros((count, MAX_N), np.uint8)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    i32p = ctypes.POINTER(ctypes.c_int32)
    _lib.check(_lib.load().tdp_canonical_batch(count, S.ctypes.data_as(u8p), P.ctypes.data_as(u8p),
                                               ns.ctypes.data_as(i32p),
                                               sig.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               h.ctypes.data_as(i32p), idx.ctypes.data_as(i32p),
                                               out.ctypes.data_as(u8p)))
    return [Canonical(sig[c, :ns[c]].copy(), int(h[c]), int(idx[c]), [int(x) for x in out[c, :ns[c]]])
            for c in range(count)]


def canonical_rows(spis: np.ndarray, ns: np.ndarray):
    """canonical_batch on pre-packed rows: spis (count, 64) uint8 one-line
    spi (entries past ns[c] ignored), pi the identity.  Returns (signatures
    float32 (count, 64), hashes int32, indices int32, canonical spis uint8
    (count, 64), entries past ns[c] zero)."""
    count = len(ns)
    S = np.ascontiguousarray(spis, np.uint8)
    if S.shape != (count, MAX_N):
        raise ValueError("spis must be (count, 64) uint8")
    nsa = np.ascontiguousarray(ns, np.int32)
    P = np.ascontiguousarray(np.broadcast_to(np.arange(MAX_N, dtype=np.uint8), (count, MAX_N)))
    sig = np.zeros((count, MAX_N), np.float32)
    h = np.zeros(count, np.int32)
    idx = np.zeros(count, np.int32)
    out = np.zeros((count, MAX_N), np.uint8)
    if count == 0:
        return sig, h, idx, out
    u8p = ctypes.POINTER(ctypes.c_uint8)
    i32p = ctypes.POINTER(ctypes.c_int32)
    _lib.check(_lib.load().tdp_canonical_batch(count, S.ctypes.data_as(u8p), P.ctypes.data_as(u8p),
                                               nsa.ctypes.data_as(i32p),
                                               sig.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               h.ctypes.data_as(i32p), idx.ctypes.data_as(i32p),
                                               out.ctypes.data_as(u8p)))
    return sig, h, idx, out


def canonical(spi: str) -> Canonical:
    """Canonical form of one configuration given as a cycle string over the
    canonical pi (new Configuration(spi).getCanonical())."""
    cycles = parse_cycles(spi)
    n = sum(len(c) for c in cycles)
    return canonical_batch([one_line(cycles, n)])[0]


def main(argv=None):
    """Times tdp_canonical_batch on seeded fixed-point-free configurations:
    python -m tdp1375_amd.canonical --count 4096 --n 9:36"""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description="batched canonicalisation timing")
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--n", default="9:36", help="lo:hi symbols per configuration")
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args(argv)
    lo, hi = (int(x) for x in args.n.split(":"))
    rng = np.random.default_rng(args.seed)
    spis = []
    while len(spis) < args.count:
        n = int(rng.integers(lo, hi + 1))
        p = rng.permutation(n)
        if n > 1 and not np.any(p == np.arange(n)):
            spis.append([int(x) for x in p])
    canonical_batch(spis[:64])   # warm-up
    times = []
    for _ in range(args.repeat):
        t0 = time.perf_counter()
        out = canonical_batch(spis)
        times.append(time.perf_counter() - t0)
    best = min(times)
    print(json.dumps({"workload": f"canonicalise {args.count} configurations, n {lo}..{hi}", "count": args.count,
                      "seconds_per_call": best, "configurations_per_s": args.count / best,
                      "distinct_hashes": len({r.hash for r in out})}), flush=True)


if __name__ == "__main__":
    main()
