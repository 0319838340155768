"""CPU restatement of the extension generators (row f3).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
tdp1375_amd.extensions; the product package never imports it.

Follows the reference step by step, one configuration and one extension at a
time, in float32 where Java computes in float and float64 where it computes
in double:
  * type 1 / 2 / 3 extensions     Extensions.java:95-203
    (unorientedExtension :281-290, remainsUnoriented :255-270, extend
    :230-253, areSymbolsInCyclicOrder(float[], float[]) :205-228,
    isOriented :272-279)
  * combinations                  Combinations.java:64-94
  * open gates                    CommonOperations.getOpenGates
  * signature / ofSignature       Configuration.java:53-154

Pinned in tests/test_extensions.py by the configurations the reference's
proof pages name (index.html:44-59), the closed-form counts of type 1, type 2
and combination extensions, the positions of the inserted symbols, and the
2-move criterion of Extensions.java:192 checked by brute force.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Set, Tuple

import numpy as np

F32 = np.float32
F64 = np.float64
Cycles = List[List[int]]


def parse(text: str) -> Cycles:
    """'(0,4,2)(1,5,3)' or '(0 4 2)(1 5 3)' -> [[0, 4, 2], [1, 5, 3]]."""
    out = []
    for part in text.replace(",", " ").split(")"):
        part = part.strip().lstrip("(")
        if part:
            out.append([int(x) for x in part.split()])
    return out


# Configuration.java roots of the two searches.
ORIENTED_5_CYCLE = parse("(0,3,1,4,2)")
INTERLEAVING_PAIR = parse("(0,4,2)(1,5,3)")
INTERSECTING_PAIR = parse("(0,3,1)(2,5,4)")
NECKLACE_SIZE_4 = parse("(0,10,2)(1,5,3)(4,8,6)(7,11,9)")
TWISTED_NECKLACE_SIZE_4 = parse("(0,7,5)(1,11,9)(2,6,4)(3,10,8)")
NECKLACE_SIZE_5 = parse("(0,4,2)(1,14,12)(3,7,5)(6,10,8)(9,13,11)")
NECKLACE_SIZE_6 = parse("(0,16,2)(1,5,3)(4,8,6)(7,11,9)(10,14,12)(13,17,15)")

# Combinations.java:44-46, with the info strings of :77-88
BAD_SMALL_COMPONENTS = [
    (ORIENTED_5_CYCLE, "bad oriented 5-cycle"),
    (INTERLEAVING_PAIR, "unoriented interleaving pair"),
    (NECKLACE_SIZE_4, "unoriented necklace of size 4"),
    (TWISTED_NECKLACE_SIZE_4, "unoriented twisted necklace of size 4"),
    (NECKLACE_SIZE_5, "unoriented necklace of size 5"),
    (NECKLACE_SIZE_6, "unoriented necklace of size 6"),
]
# Extensions.java:41-46
DFS_ROOTS = [ORIENTED_5_CYCLE, INTERLEAVING_PAIR, INTERSECTING_PAIR]


def size(cycles: Cycles) -> int:
    return sum(len(c) for c in cycles)


def _cyclic(idx, symbols) -> bool:
    """CommonOperations.areSymbolsInCyclicOrder (CommonOperations.java:167-182):
    at most one descent of idx along the symbols, taken cyclically."""
    leap = False
    m = len(symbols)
    for t in range(m):
        if idx[symbols[t]] > idx[symbols[(t + 1) % m]]:
            if leap:
                return False
            leap = True
    return True


def signature(cycles: Cycles, n: Optional[int] = None) -> np.ndarray:
    """Configuration.signature with pi the identity (Configuration.java:53-101):
    the label of each position's cycle by first appearance, plus rank/100 for
    an oriented cycle (ranks along the cycle from its first-positioned symbol)."""
    n = size(cycles) if n is None else n
    of = [0] * n
    for ci, c in enumerate(cycles):
        for s in c:
            of[s] = ci
    inv_idx = [(n - x) % n for x in range(n)]     # index of x in pi^-1
    oriented = [not _cyclic(inv_idx, c) for c in cycles]
    rank: Dict[int, int] = {}
    for ci, c in enumerate(cycles):
        if oriented[ci]:
            m = c.index(min(c))
            for k in range(len(c)):
                rank[c[(m + k) % len(c)]] = k + 1
    label = {}
    sig = np.zeros(n, F32)
    for i in range(n):
        ci = of[i]
        if ci not in label:
            label[ci] = F32(len(label) + 1)
        v = label[ci]
        if oriented[ci]:
            v = F32(v + F32(F32(rank[i]) / F32(100)))
        sig[i] = v
    return sig


def of_signature(sig: Sequence[float]) -> Cycles:
    """Configuration.ofSignature (Configuration.java:102-154): per label, the
    positions from last to first, or the positions in rank order when the
    label has fractional entries; cycles in label order."""
    sig = np.asarray(sig, F32)
    n = len(sig)
    labels = [int(v) for v in sig]
    maxl = max(labels) if n else 0
    plain: List[List[int]] = [[] for _ in range(maxl + 1)]
    ranked: List[Dict[int, int]] = [dict() for _ in range(maxl + 1)]
    for i in range(n - 1, -1, -1):
        v, l = sig[i], labels[i]
        plain[l].append(i)
        if np.fmod(v, F32(1)) > 0:
            r = int(np.floor(F32(F32(v - F32(l)) * F32(100)) + F32(0.5)))
            ranked[l][r] = i
    out = []
    for l in range(maxl + 1):
        if not plain[l]:
            continue
        out.append([ranked[l][r] for r in sorted(ranked[l])] if ranked[l] else plain[l])
    return out


def open_gates(cycles: Cycles, n: Optional[int] = None) -> Set[int]:
    """CommonOperations.getOpenGates with pi the identity: a of (a, spi(a)) is
    an open gate unless another cycle has a pair (c, d) with a, c, b, d in the
    cyclic order of pi^-1, or its own cycle has c with a, b, c in the cyclic
    order of pi."""
    n = size(cycles) if n is None else n
    inv_idx = [(n - x) % n for x in range(n)]
    idx = list(range(n))
    gates = set()
    nontrivial = [c for c in cycles if len(c) > 1]
    for ei, e in enumerate(nontrivial):
        for t in range(len(e)):
            a, b = e[t], e[(t + 1) % len(e)]
            closed = False
            for gi, g in enumerate(nontrivial):
                if gi != ei:
                    for u in range(len(g)):
                        if _cyclic(inv_idx, (a, g[u], b, g[(u + 1) % len(g)])):
                            closed = True
                            break
                else:
                    for c in e:
                        if c != a and c != b and _cyclic(idx, (a, b, c)):
                            closed = True
                            break
                if closed:
                    break
            if not closed:
                gates.add(a)
    return gates


def is_full(cycles: Cycles, n: Optional[int] = None) -> bool:
    return not open_gates(cycles, n)


def _insert(sig: np.ndarray, label, positions) -> np.ndarray:
    """Extensions.unorientedExtension (Extensions.java:281-290)."""
    out = list(sig)
    for t, p in enumerate(sorted(positions)):
        out.insert(p + t, F32(label))
    return np.asarray(out, F32)


def type1_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """Extensions.java:95-116."""
    sig = signature(cycles)
    n = len(sig)
    new = len(cycles) + 1
    gates = open_gates(cycles, n)
    out = []
    for i in range(n):
        if i not in gates:
            continue
        for b in range(n):
            for c in range(b, n):
                if not (i == b == c):
                    out.append((f"a={i} b={b} c={c}", of_signature(_insert(sig, new, (i, b, c)))))
    return out


def type2_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """Extensions.java:121-144."""
    sig = signature(cycles)
    n = len(sig)
    if open_gates(cycles, n):
        return []
    new = len(cycles) + 1
    out = []
    for a in range(n):
        for b in range(a, n):
            for c in range(b, n):
                if not (a == b == c):
                    out.append((f"a={a} b={b} c={c}", of_signature(_insert(sig, new, (a, b, c)))))
    return out


def _remains_unoriented(indexes: List[int], *new) -> bool:
    """Extensions.remainsUnoriented (:255-270): every new index falls in the
    same gap between consecutive positions of the cycle."""
    gaps = set()
    m = len(indexes)
    for x in new:
        for t in range(m):
            left, right = indexes[t], indexes[(t + 1) % m]
            if (left < x <= right) or (right < left and (left < x or x <= right)):
                gaps.add((left, right))
    return len(gaps) == 1


def _turn_oriented(cycle: List[int], label: int, sig: np.ndarray, a: int, b: int) -> np.ndarray:
    """Extensions.extend (:230-253): the 3-cycle's symbols from its largest get
    label + 0.01, 0.02, 0.03 (computed in double, stored as float), then
    label + 0.05 is inserted before position a and label + 0.04 before b."""
    m = cycle.index(max(cycle))
    rot = cycle[m:] + cycle[:m]
    out = [F32(v) for v in sig]
    for t in range(len(rot) - 1, -1, -1):
        out[rot[t]] = F32(F64(out[rot[t]]) + 0.01 * (t + 1))
    nxt = F32(0.05)
    for ins, p in enumerate((a, b)):
        out.insert(p + ins, F32(F32(label) + nxt))
        nxt = F32(nxt - F32(0.01))
    return np.asarray(out, F32)


_RANK_ORDER = (0.01, 0.03, 0.05, 0.02, 0.04)


def _in_rank_order(elements: np.ndarray, other: np.ndarray) -> bool:
    """Extensions.areSymbolsInCyclicOrder(float[], float[]) (:205-228),
    including its answer when no element equals other[0]."""
    nxt = 0
    n, m = len(elements), len(other)
    for i in range(n):
        if elements[i] == other[nxt]:
            for j in range(n + 1):
                e = elements[(i + j) % n]
                o = other[nxt % m]
                if np.floor(F64(e)) == np.floor(F64(o)):
                    if e == o:
                        nxt += 1
                        if nxt > m:
                            return True
                    else:
                        return False
    return True


def type3_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """Extensions.java:149-203."""
    sig = signature(cycles)
    n = len(sig)
    by_label: Dict[int, List[int]] = {}
    for i in range(n):
        by_label.setdefault(int(np.floor(sig[i])), []).append(i)
    cycle_of = {}
    for c in cycles:
        for s in c:
            cycle_of[s] = c
    cycle_by_label = {l: cycle_of[ps[0]] for l, ps in by_label.items()}
    out = []
    for label in range(1, len(cycles) + 1):
        if any(np.fmod(v, F32(1)) > 0 and int(np.floor(v)) == label for v in sig):
            continue
        cyc = cycle_by_label[label]
        cyc_text = "(" + " ".join(str(s) for s in _from_min(cyc)) + ")"
        fractions = np.asarray([F32(F32(f) + F32(label)) for f in _RANK_ORDER], F32)
        for a in range(n):
            for b in range(a, n):
                ext = of_signature(_insert(sig, label, (a, b)))
                if _remains_unoriented(by_label[label], a, b):
                    if len(open_gates(ext)) <= 2:
                        out.append((f"a={a} b={b}, extended cycle: {cyc_text}", ext))
                elif len(by_label[label]) == 3:
                    prime = _turn_oriented(cyc, label, sig, a, b)
                    if _in_rank_order(prime, fractions):
                        out.append((f"a={a} b={b}, extended cycle: {cyc_text}, turn oriented", of_signature(prime)))
    return out


def _from_min(c: List[int]) -> List[int]:
    m = c.index(min(c))
    return c[m:] + c[:m]


def dfs_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """SortOrExtendExtensions.extend (Extensions.java:319-324): type 1, 2, 3."""
    return type1_extensions(cycles) + type2_extensions(cycles) + type3_extensions(cycles)


def combination_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """Combinations.extend (:64-94): every bad small component, labels shifted
    by the number of cycles (float32 sums), inserted before every position."""
    sig = signature(cycles)
    k = F32(len(cycles))
    out = []
    for comp, info in BAD_SMALL_COMPONENTS:
        shifted = [F32(v + k) for v in signature(comp)]
        for i in range(len(sig)):
            ext = np.asarray(list(sig[:i]) + shifted + list(sig[i:]), F32)
            out.append((f"pos={i} add {info}", of_signature(ext)))
    return out
