"""CPU restatement of configuration canonicalisation (row f2).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
canonicalisation kernel; the product package never imports it.

Follows Configuration.java (reference):
  * signature(spi, pi)            Configuration.java:53-101
  * equivalent signatures         Configuration.java:182-195 (2n: every rotation
    of pi, each with its mirror)
  * canonical signature and hash  Configuration.java:159-180, 283-297 (least
    CRC32 of the big-endian float32 signature as a signed int, ties broken by
    Arrays.compare of the contents)
  * ofSignature                   Configuration.java:103-157
  * areSymbolsInCyclicOrder       CommonOperations.java:167-182

A permutation spi is given as a list of cycles (symbol lists) or as its
one-line image array; pi as the list of symbols in position order.  Values
are float32 as in the reference (label + (float) k / 100 in single precision).

Pinned by the golden vectors of ConfigurationTest.java (signatures, all 2n
equivalent signatures, canonical signatures and hash codes of three
configurations) in tests/golden/configuration_vectors.json, and by the
round trip ofSignature(signature) -> signature (ConfigurationTest.java
round-trip test).  zlib.crc32 is the same CRC-32 as java.util.zip.CRC32.
"""
from __future__ import annotations

import re
import struct
import zlib
from typing import List, Sequence, Tuple

import numpy as np

F32 = np.float32


def parse_cycles(text: str) -> List[List[int]]:
    """'(0 9 7)(1 5 2 6 3)' -> [[0, 9, 7], [1, 5, 2, 6, 3]] (MulticyclePermutation.java:78-88)."""
    return [[int(x) for x in re.split(r"[,\s]+", g.strip()) if x] for g in re.findall(r"\(([^()]*?)\)", text)]


def image_of(cycles: Sequence[Sequence[int]], n: int) -> List[int]:
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


def in_cyclic_order(index: Sequence[int], symbols: Sequence[int]) -> bool:
    """CommonOperations.java:167-182: at most one descent of index along the
    cyclic sequence."""
    leap = False
    for i in range(len(symbols)):
        if index[symbols[i]] > index[symbols[(i + 1) % len(symbols)]]:
            if leap:
                return False
            leap = True
    return True


def signature(cycles: Sequence[Sequence[int]], pi: Sequence[int]) -> np.ndarray:
    """Configuration.java:53-101."""
    n = len(pi)
    pos = [0] * (max(pi) + 1)
    for i, s in enumerate(pi):
        pos[s] = i
    inv_index = [0] * (max(pi) + 1)          # index of each symbol in pi^-1
    for i, s in enumerate(reversed(list(pi))):
        inv_index[s] = i
    cycle_of = {}
    for ci, c in enumerate(cycles):
        for s in c:
            cycle_of[s] = ci
    labels, oriented, sym_index = {}, {}, {}
    out = np.zeros(n, F32)
    for i in range(n):
        s = pi[i]
        ci = cycle_of[s]
        c = cycles[ci]
        if ci not in oriented:
            oriented[ci] = not in_cyclic_order(inv_index, c)
        if oriented[ci] and ci not in sym_index:
            m = min(range(len(c)), key=lambda j: pos[c[j]])
            sym_index[ci] = {c[(m + k) % len(c)]: k + 1 for k in range(len(c))}
        if ci not in labels:
            labels[ci] = len(labels) + 1
        label = F32(labels[ci])
        out[i] = label + F32(sym_index[ci][s]) / F32(100) if oriented[ci] else label
    return out


def mirror_cycles(cycles: Sequence[Sequence[int]], pi: Sequence[int]) -> List[List[int]]:
    """Configuration.java:191-198: spi conjugated by prod (pi[i] pi[n-1-i]),
    then inverted (each cycle relabelled, then reversed)."""
    n = len(pi)
    tau = {s: s for s in pi}
    for i in range(n // 2):
        tau[pi[i]], tau[pi[n - 1 - i]] = pi[n - 1 - i], pi[i]
    return [[tau[s] for s in reversed(c)] for c in cycles]


def equivalent_signatures(cycles: Sequence[Sequence[int]], pi: Sequence[int]) -> List[Tuple[np.ndarray, bool, int]]:
    """Configuration.java:182-189: for every rotation r of pi, its signature
    and its mirror's (in that order)."""
    out = []
    for r in range(len(pi)):
        pr = list(pi[r:]) + list(pi[:r])
        out.append((signature(cycles, pr), False, r))
        out.append((signature(mirror_cycles(cycles, pr), pr), True, r))
    return out


def crc_hash(sig: np.ndarray) -> int:
    """Configuration.java:283-297: CRC32 of the big-endian float32 bytes, as a
    signed 32-bit int."""
    h = zlib.crc32(np.asarray(sig, F32).astype(">f4").tobytes()) & 0xFFFFFFFF
    return h - (1 << 32) if h >= (1 << 31) else h


def canonical(cycles: Sequence[Sequence[int]], pi: Sequence[int]):
    """Configuration.java:159-176: (signature, index in the equivalent list)
    with the least hash, ties broken by the lexicographically least content."""
    best, best_h, best_i = None, None, -1
    for idx, (sig, _, _) in enumerate(equivalent_signatures(cycles, pi)):
        h = crc_hash(sig)
        if best is None or h < best_h or (h == best_h and list(sig) < list(best)):
            best, best_h, best_i = sig, h, idx
    return best, best_h, best_i


def of_signature(sig: Sequence[float]) -> List[List[int]]:
    """Configuration.java:103-157 (pi canonical: the symbol at position i is i)."""
    sig = np.asarray(sig, F32)
    n = len(sig)
    labels = [int(v) for v in sig]
    maxl = max(labels)
    members = {l: [] for l in range(maxl + 1)}
    oriented = {l: {} for l in range(maxl + 1)}
    for i in range(n - 1, -1, -1):
        v, l = sig[i], labels[i]
        members[l].append(i)
        if float(v) % 1 > 0:
            rank = int(np.floor((v - F32(l)) * F32(100) + F32(0.5)))
            oriented[l][rank] = i
    cycles = []
    for l in range(maxl + 1):
        if not members[l]:
            continue
        if oriented[l]:
            cycles.append([oriented[l][r] for r in sorted(oriented[l])])
        else:
            cycles.append(members[l])
    return cycles


def config_hash(cycles: Sequence[Sequence[int]], pi: Sequence[int]) -> int:
    """Configuration.hashCode (Configuration.java:243-247): the hash of the
    canonical configuration's own signature."""
    sig, _, _ = canonical(cycles, pi)
    canon = of_signature(sig)
    return crc_hash(signature(canon, list(range(len(pi)))))
