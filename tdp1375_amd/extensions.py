"""Extension generators: the producers of search batches, and a batched
level-by-level SortOrExtend exploration over them.

Generators (host code, as in the reference; exact float32 arithmetic where
the reference computes in float, float64 where it computes in double):

* ``type1_extensions`` / ``type2_extensions`` / ``type3_extensions``
  (Extensions.java:95-203): a new unoriented 3-cycle with one symbol in an
  open gate; a new 3-cycle anywhere in a full configuration; an unoriented
  cycle grown by two symbols, either staying unoriented (at most two open
  gates after) or, for a 3-cycle, turned into an oriented 5-cycle whose
  symbols are in the cyclic order 1,3,5,2,4 of its ranks.
* ``combination_extensions`` (Combinations.java:64-94,183-199): a bad small
  component inserted at every position of the signature.

Each returns ``(info, cycles)`` pairs in the reference's order, with its info
string; configurations have pi the identity (``Configuration.ofSignature``).

``explore`` runs the SortOrExtend tree (SortOrExtend.java:51-87 with the
extend overrides Extensions.java:313-325 and Combinations.java:168-181) one
level at a time instead of one ForkJoin task per configuration: every level's
configurations are canonicalised in one GPU launch (tdp_canonical_batch), the
canonical forms not seen before are searched with the batched depth ladder
(one persistent launch per rung), and every configuration whose canonical
form has no sorting is extended into the next level.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field
from itertools import compress
from multiprocessing import Pool
from typing import Callable, Dict, List, Optional, Sequence, Set, Tuple

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


def signature(cycles: Cycles, n: Optional[int] = None) -> np.ndarray:
    """Configuration.signature with pi the identity (Configuration.java:53-101):
    the label of each position's cycle by first appearance, plus rank/100 in
    float32 for an oriented cycle (ranks along the cycle from its smallest
    symbol, the first of its symbols along pi)."""
    n = size(cycles) if n is None else n
    p, cid = _one_line(cycles, n)
    sig = np.zeros(n, F32)
    label: Dict[int, int] = {}
    for s in range(n):
        label.setdefault(int(cid[s]), len(label) + 1)
    inv = (n - np.arange(n)) % n                         # index of each symbol in pi^-1
    for k, c in enumerate(cycles):
        lab = F32(label[k])
        ic = inv[c]
        if int(np.count_nonzero(ic > np.roll(ic, -1))) <= 1:   # in the cyclic order of pi^-1
            sig[c] = lab
        else:
            m = c.index(min(c))
            for r in range(len(c)):
                sig[c[(m + r) % len(c)]] = F32(lab + F32(F32(r + 1) / F32(100)))
    return sig


def _one_line(cycles: Cycles, n: int):
    p = np.arange(n)
    cid = np.full(n, -1)
    for k, c in enumerate(cycles):
        p[c] = np.roll(c, -1)
        cid[c] = k
    return p, cid


def of_signature(sig: Sequence[float]) -> Cycles:
    """Configuration.ofSignature (Configuration.java:102-154): per label, the
    positions from last to first, or, when the label has fractional entries,
    those positions in the order of their ranks round((v - label) * 100) in
    float32; cycles in label order."""
    sig = np.asarray(sig, F32)
    n = len(sig)
    lab = sig.astype(np.int64)
    frac = np.fmod(sig, F32(1)) > 0
    rank = np.floor((sig - lab.astype(F32)) * F32(100) + F32(0.5)).astype(np.int64)
    order = np.lexsort((-np.arange(n), lab))
    cut = np.flatnonzero(np.diff(lab[order])) + 1
    out = []
    for grp in np.split(order, cut):
        f = frac[grp]
        if f.any():
            g = grp[f]
            out.append(g[np.argsort(rank[g], kind="stable")].tolist())
        else:
            out.append(grp.tolist())
    return out


def open_gates(cycles: Cycles, n: Optional[int] = None) -> Set[int]:
    """CommonOperations.getOpenGates with pi the identity, for all pairs at
    once: a of the pair (a, spi(a)) is an open gate unless a pair (c, spi(c))
    of another cycle has a, c, spi(a), spi(c) in the cyclic order of pi^-1,
    or a symbol c of its own cycle has a, spi(a), c in the cyclic order of pi
    (at most one descent, taken cyclically)."""
    n = size(cycles) if n is None else n
    p = np.arange(n)
    cid = np.full(n, -1)
    for k, c in enumerate(cycles):
        if len(c) > 1:
            p[c] = np.roll(c, -1)
            cid[c] = k
    A = np.flatnonzero(cid >= 0)
    if A.size == 0:
        return set()
    B = p[A]
    inv = (n - np.arange(n)) % n
    ia, ib = inv[A][:, None], inv[B][:, None]
    ic, idd = inv[A][None, :], inv[B][None, :]
    desc4 = (ia > ic).astype(np.int8) + (ic > ib) + (ib > idd) + (idd > ia)
    same = cid[A][:, None] == cid[A][None, :]
    crossing = ((desc4 <= 1) & ~same).any(axis=1)
    a_, b_, c_ = A[:, None], B[:, None], A[None, :]
    desc3 = (a_ > b_).astype(np.int8) + (b_ > c_) + (c_ > a_)
    triple = ((desc3 <= 1) & same & (c_ != a_) & (c_ != b_)).any(axis=1)
    return set(A[~crossing & ~triple].tolist())


def is_full(cycles: Cycles, n: Optional[int] = None) -> bool:
    return not open_gates(cycles, n)


def _split(rows: List[List[int]], lens: List[int]) -> List[Cycles]:
    offs = np.concatenate([[0], np.cumsum(lens)]).tolist()
    spans = list(zip(offs[:-1], offs[1:]))
    return [[r[i:j] for i, j in spans] for r in rows]


def _new_3cycle_extensions(cycles: Cycles, infos: List[str], triples: List[Tuple[int, int, int]]
                           ) -> List[Tuple[str, Cycles]]:
    """Insertion of a new label (the largest) before sorted original positions
    p0 <= p1 <= p2 of the signature (Extensions.unorientedExtension,
    :281-290), for all triples at once: ofSignature gives the parent's cycles
    in label order with every symbol moved past the insertions before it, then
    the new cycle, its positions from last to first."""
    if not triples:
        return []
    sig = signature(cycles)
    n = len(sig)
    base = of_signature(sig)
    flat = np.concatenate([np.asarray(c) for c in base])
    T = np.asarray(triples)
    ar = np.arange(n)[None, :]
    M = ar + (ar >= T[:, 0:1]) + (ar >= T[:, 1:2]) + (ar >= T[:, 2:3])
    rows = _split(M[:, flat].tolist(), [len(c) for c in base])
    return [(info, ext + [[p2 + 2, p1 + 1, p0]]) for info, ext, (p0, p1, p2) in zip(infos, rows, triples)]


def type1_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """Extensions.java:95-116: for every open gate i and b <= c, a new 3-cycle
    before positions i, b, c (not all equal)."""
    n = size(cycles)
    gates = open_gates(cycles, n)
    infos, triples = [], []
    for i in range(n):
        if i in gates:
            for b in range(n):
                for c in range(b, n):
                    if not (i == b == c):
                        infos.append(f"a={i} b={b} c={c}")
                        triples.append(tuple(sorted((i, b, c))))
    return _new_3cycle_extensions(cycles, infos, triples)


def type2_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """Extensions.java:121-144: a full configuration gets a new 3-cycle before
    every a <= b <= c (not all equal)."""
    n = size(cycles)
    if open_gates(cycles, n):
        return []
    infos, triples = [], []
    for a in range(n):
        for b in range(a, n):
            for c in range(b, n):
                if not (a == b == c):
                    infos.append(f"a={a} b={b} c={c}")
                    triples.append((a, b, c))
    return _new_3cycle_extensions(cycles, infos, triples)


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
    out = sig.copy()
    for t in range(len(rot) - 1, -1, -1):
        out[rot[t]] = F32(F64(out[rot[t]]) + 0.01 * (t + 1))
    first = F32(F32(label) + F32(0.05))
    second = F32(F32(label) + F32(F32(0.05) - F32(0.01)))
    return np.insert(out, [a, b], [first, second]).astype(F32)


_RANK_ORDER = (0.01, 0.03, 0.05, 0.02, 0.04)


def _in_rank_order(elements: np.ndarray, other: np.ndarray) -> bool:
    """Extensions.areSymbolsInCyclicOrder(float[], float[]) (:205-228): from
    the first element equal to other[0], once around the circle and back to
    it, every element with other's label must equal other[0], other[1], ...
    in turn (the grown cycle has as many elements as other, so the first such
    element decides); true as well when no element equals other[0]."""
    hit = np.flatnonzero(elements == other[0])
    if hit.size == 0:
        return True
    ring = np.roll(elements, -int(hit[0])).tolist()
    oth = other.tolist()
    m = len(oth)
    nxt = 0
    for e in ring + ring[:1]:
        o = oth[nxt % m]
        if math.floor(e) == math.floor(o):
            if e != o:
                return False
            nxt += 1
            if nxt > m:
                return True
    return True


def _open_gate_counts(perms: np.ndarray, cid: np.ndarray) -> np.ndarray:
    """Number of open gates (as in open_gates) of each of T configurations of
    N symbols at once: perms (T, N) one-line spi, cid (T, N) cycle index of
    each symbol (-1 for a fixed point)."""
    T, N = perms.shape
    inv = (N - np.arange(N)) % N
    ia = inv[None, :, None]
    ib = inv[perms][:, :, None]
    ic = inv[None, None, :]
    idd = inv[perms][:, None, :]
    desc4 = (ia > ic).astype(np.int8) + (ic > ib) + (ib > idd) + (idd > ia)
    same = cid[:, :, None] == cid[:, None, :]
    live = cid >= 0
    pair = live[:, :, None] & live[:, None, :]
    crossing = ((desc4 <= 1) & ~same & pair).any(axis=2)
    a_ = np.arange(N)[None, :, None]
    b_ = perms[:, :, None]
    c_ = np.arange(N)[None, None, :]
    desc3 = (a_ > b_).astype(np.int8) + (b_ > c_) + (c_ > a_)
    triple = ((desc3 <= 1) & same & (c_ != a_) & (c_ != b_) & pair).any(axis=2)
    return (live & ~crossing & ~triple).sum(axis=1)


def type3_extensions(cycles: Cycles) -> List[Tuple[str, Cycles]]:
    """Extensions.java:149-203: an unoriented cycle grown by two symbols
    before positions a <= b; kept when it stays unoriented and leaves at most
    two open gates, or, for a 3-cycle that turns oriented, when its ranks lie
    in the cyclic order 1, 3, 5, 2, 4 along pi (otherwise the configuration
    accepts a 2-move).

    The turned cycle's values are label + 0.01, 0.02, 0.03 (float of a double
    sum) on its symbols from the largest, and label + 0.05, label + 0.04
    (float sums) inserted before a and b; for every label up to 199 these
    equal the float sums label + 0.01f ... 0.05f they are compared with, so
    the comparison of Extensions.java:205-228 is the cyclic order of those
    five positions, and ofSignature orders the cycle by rank: the three
    symbols from the largest, then b + 1, then a."""
    sig = signature(cycles)
    n = len(sig)
    lab = sig.astype(np.int64)
    frac = np.fmod(sig, F32(1)) > 0
    base = of_signature(sig)
    cycle_of = {}
    for c in cycles:
        for s in c:
            cycle_of[s] = c
    entries = []          # (info, ext, needs the open-gate test)
    for label in range(1, len(cycles) + 1):
        where = np.flatnonzero(lab == label)
        if frac[where].any():
            continue
        pos = where.tolist()
        cyc = cycle_of[pos[0]]
        m0 = cyc.index(min(cyc))
        cyc_text = "(" + " ".join(str(s) for s in cyc[m0:] + cyc[:m0]) + ")"
        k = label - 1                                   # its index in base (label order)
        mx = cyc.index(max(cyc))
        rot = cyc[mx:] + cyc[:mx]
        for a in range(n):
            for b in range(a, n):
                stays = _remains_unoriented(pos, a, b)
                if not stays and len(pos) != 3:
                    continue
                shift = lambda x: x + (x >= a) + (x >= b)
                ext = [[shift(x) for x in c] for c in base]
                if stays:
                    ext[k] = sorted(ext[k] + [a, b + 1], reverse=True)
                    entries.append((f"a={a} b={b}, extended cycle: {cyc_text}", ext, True))
                else:
                    q0, q1, q2 = (shift(x) for x in rot)
                    order = [q0, q2, a, q1, b + 1]
                    d = [(x - q0) % (n + 2) for x in order]
                    if d[1] < d[2] < d[3] < d[4]:
                        ext[k] = [q0, q1, q2, b + 1, a]
                        entries.append((f"a={a} b={b}, extended cycle: {cyc_text}, turn oriented", ext, False))
    test = [e for e in entries if e[2]]
    keep = {}
    if test:
        N = n + 2
        rows_p, rows_c = [], []
        for _, ext, _ in test:
            perm = list(range(N))
            cid = [-1] * N
            for ci, c in enumerate(ext):
                for t in range(len(c)):
                    perm[c[t]] = c[t + 1 - len(c)]
                    cid[c[t]] = ci
            rows_p.append(perm)
            rows_c.append(cid)
        counts = _open_gate_counts(np.asarray(rows_p), np.asarray(rows_c))
        keep = {id(e[1]): int(g) <= 2 for e, g in zip(test, counts)}
    return [(info, ext) for info, ext, gated in entries if not gated or keep[id(ext)]]


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
        shifted = (signature(comp) + k).astype(F32)
        for i in range(len(sig)):
            out.append((f"pos={i} add {info}", of_signature(np.concatenate([sig[:i], shifted, sig[i:]]))))
    return out


# ---------------------------------------------------------------- one-line rows
# The same extensions as one-line spi rows (pi the identity), built for all
# insertion positions of a parent at once; explore works on these.  A result
# is (rows, ns): rows (K, 64) uint8 padded with 255, ns (K,) the sizes; row r
# is the permutation of the r-th configuration the matching (info, cycles)
# generator above returns.
ROW_WIDTH = 64
PAD = 255


def _pad(parts: List[np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    parts = [q for q in parts if len(q)]
    K = sum(len(q) for q in parts)
    rows = np.full((K, ROW_WIDTH), PAD, np.uint8)
    ns = np.empty(K, np.int32)
    at = 0
    for q in parts:
        if q.shape[1] > ROW_WIDTH:
            raise ValueError("an extension exceeds 64 symbols")
        rows[at:at + len(q), :q.shape[1]] = q
        ns[at:at + len(q)] = q.shape[1]
        at += len(q)
    return rows, ns

def cycles_of_row(row: Sequence[int]) -> Cycles:
    """Cycles of a one-line spi, each from its smallest symbol, in order of
    their smallest symbols (the order of their labels in the signature)."""
    p = [int(x) for x in row]
    seen = [False] * len(p)
    out = []
    for s in range(len(p)):
        if not seen[s]:
            c, x = [], s
            while not seen[x]:
                seen[x] = True
                c.append(x)
                x = p[x]
            out.append(c)
    return out


def _moved(p: np.ndarray, M: np.ndarray, extra: int) -> np.ndarray:
    """Rows of n + extra symbols with every parent symbol s moved to M[r, s]
    (and its image to M[r, p[s]]); the inserted symbols are left to the
    caller."""
    K, n = M.shape
    out = np.empty((K, n + extra), np.int64)
    out[np.arange(K)[:, None], M] = M[:, p]
    return out


def _new_3cycle_rows(p: np.ndarray, T: np.ndarray) -> np.ndarray:
    n = len(p)
    if len(T) == 0:
        return np.empty((0, n + 3), np.int64)
    ar = np.arange(n)[None, :]
    M = ar + (ar >= T[:, 0:1]) + (ar >= T[:, 1:2]) + (ar >= T[:, 2:3])
    out = _moved(p, M, 3)
    r = np.arange(len(T))
    c0, c1, c2 = T[:, 2] + 2, T[:, 1] + 1, T[:, 0]
    out[r, c0] = c1
    out[r, c1] = c2
    out[r, c2] = c0
    return out


def _oriented(cycle: List[int], n: int) -> bool:
    inv = [(n - x) % n for x in cycle]
    return sum(1 for t in range(len(inv)) if inv[t] > inv[(t + 1) % len(inv)]) > 1


def dfs_extension_rows(row: Sequence[int]) -> Tuple[np.ndarray, np.ndarray]:
    """dfs_extensions as one-line rows (type 1, then type 2, then type 3)."""
    p = np.asarray(row, np.int64)
    n = len(p)
    cycles = cycles_of_row(p)
    gates = sorted(open_gates(cycles, n))
    bi, ci = np.triu_indices(n)
    parts = []
    # type 1: gate i, b <= c, not all equal, positions sorted
    if gates:
        I = np.repeat(np.asarray(gates), len(bi))
        B = np.tile(bi, len(gates))
        C = np.tile(ci, len(gates))
        keep = ~((I == B) & (B == C))
        parts.append(_new_3cycle_rows(p, np.sort(np.stack([I, B, C], 1)[keep], axis=1)))
    else:
        # type 2: a full configuration, a <= b <= c, not all equal
        T = np.asarray([(a, b, c) for a in range(n) for b in range(a, n) for c in range(b, n)
                        if not (a == b == c)], np.int64).reshape(-1, 3)
        parts.append(_new_3cycle_rows(p, T))
    # type 3
    A, Bp = np.triu_indices(n)
    K = len(A)
    r = np.arange(K)
    ar = np.arange(n)[None, :]
    M = ar + (ar >= A[:, None]) + (ar >= Bp[:, None])
    moved = _moved(p, M, 2)
    cid = np.empty(n, np.int64)
    for k, c in enumerate(cycles):
        cid[c] = k
    cid_moved = np.empty((K, n + 2), np.int64)
    cid_moved[r[:, None], M] = cid[None, :]
    for k, c in enumerate(cycles):                # label k + 1: cycles in order of their smallest symbols
        if _oriented(c, n):
            continue
        pos = np.asarray(sorted(c))
        L = len(pos)
        stays = (np.searchsorted(pos, A) % L) == (np.searchsorted(pos, Bp) % L)
        keep = np.zeros(K, bool)
        rows = moved.copy()
        if stays.any():
            srt = -np.sort(-np.concatenate([M[:, pos], A[:, None], Bp[:, None] + 1], axis=1), axis=1)
            grown = rows[stays]
            g = srt[stays]
            gr = np.arange(len(g))[:, None]
            grown[gr, g] = np.roll(g, -1, axis=1)
            cm = cid_moved[stays].copy()
            cm[np.arange(len(g)), A[stays]] = k
            cm[np.arange(len(g)), Bp[stays] + 1] = k
            rows[stays] = grown
            keep[stays] = _open_gate_counts(grown, cm) <= 2
        if L == 3:
            m = int(np.argmax(c))
            top = c[m]
            rot = [top, int(p[top]), int(p[p[top]])]
            turn = ~stays
            q0, q1, q2 = (M[turn][:, x] for x in rot)
            a, b1 = A[turn], Bp[turn] + 1
            d = [(x - q0) % (n + 2) for x in (q2, a, q1, b1)]
            ok = (d[0] < d[1]) & (d[1] < d[2]) & (d[2] < d[3])
            tr = rows[turn]
            t = np.arange(len(tr))
            tr[t, q0] = q1
            tr[t, q1] = q2
            tr[t, q2] = b1
            tr[t, b1] = a
            tr[t, a] = q0
            rows[turn] = tr
            keep[turn] = ok
        parts.append(rows[keep])
    return _pad(parts)


def combination_extension_rows(row: Sequence[int]) -> Tuple[np.ndarray, np.ndarray]:
    """combination_extensions as one-line rows."""
    p = np.asarray(row, np.int64)
    n = len(p)
    ar = np.arange(n)[None, :]
    I = np.arange(n)[:, None]
    parts = []
    for comp, _ in BAD_SMALL_COMPONENTS:
        pb = np.asarray(one_line_of(comp), np.int64)
        nb = len(pb)
        M = ar + nb * (ar >= I)
        out = _moved(p, M, nb)
        out[np.arange(n)[:, None], I + np.arange(nb)[None, :]] = I + pb[None, :]
        parts.append(out)
    return _pad(parts)


def one_line_of(cycles: Cycles) -> List[int]:
    n = size(cycles)
    p = list(range(n))
    for c in cycles:
        for t in range(len(c)):
            p[c[t]] = c[t + 1 - len(c)]
    return p


# ---------------------------------------------------------------- exploration
SEARCH_BATCH_MAX = 65536      # configurations per batched search (16-bit index in the dedup tag)

@dataclass
class Level:
    depth: int
    configurations: int          # distinct configurations of the level
    new_canonical: int           # canonical forms searched at this level
    found: int                   # of those, with a sorting
    extended: int                # configurations with no sorting (extended unless the level is the last)
    max_three_norm: int
    max_bad_three_norm: int      # largest 3-norm among the configurations with no sorting (-1: none)
    seconds: Dict[str, float] = field(default_factory=dict)


@dataclass
class Exploration:
    levels: List[Level]
    verdicts: Dict[Tuple[int, ...], bool]             # canonical one-line spi -> sorting found
    sortings: Dict[Tuple[int, ...], List[Tuple[int, int, int]]]
    truncated: bool


def _keys(rows: np.ndarray) -> List[bytes]:
    return np.ascontiguousarray(rows).view(np.dtype((np.void, ROW_WIDTH))).ravel().tolist()


def _extend_chunk(args) -> List[Tuple[np.ndarray, np.ndarray]]:
    extend_rows, rows = args
    return [extend_rows(r) for r in rows]


def explore(roots: Sequence[Cycles], extend_rows: Callable[[Sequence[int]], Tuple[np.ndarray, np.ndarray]],
            max_levels: int, max_ratio: float = 1.375, max_depth: int = 11, slot: int = 0,
            max_configurations: int = 1 << 21, log: Optional[Callable[[str], None]] = None,
            workers: int = 1) -> Exploration:
    """The SortOrExtend tree from roots, level by level, to max_levels levels,
    with extend_rows (dfs_extension_rows or combination_extension_rows).

    A configuration already visited is not visited again (its extensions
    would be the same); a canonical form is searched once, the first time it
    is reached (the single-flight coordinator and the sorting / bad-case files
    of the reference); every configuration whose canonical form has no
    sorting is extended.  A level with more than max_configurations distinct
    configurations stops the exploration (truncated=True).  verdicts and
    sortings are keyed by the canonical one-line spi as a tuple.  With
    workers > 1 the configurations of a level are extended on that many host
    processes (extend_rows must be a module-level function), in chunks whose
    results are joined in the same order as one process would."""
    from . import canonical as K
    from . import proof_driver as D

    seen: Set[bytes] = set()
    verdicts: Dict[bytes, bool] = {}
    norms: Dict[bytes, int] = {}
    sortings: Dict[Tuple[int, ...], List[Tuple[int, int, int]]] = {}
    levels: List[Level] = []
    rows, ns = _pad([np.asarray([one_line_of(r)], np.int64) for r in roots])
    truncated = False
    pool = None
    try:
        for depth in range(max_levels):
            t0 = time.perf_counter()
            view = np.ascontiguousarray(rows).view(np.dtype((np.void, ROW_WIDTH))).ravel()
            first = np.sort(np.unique(view, return_index=True)[1])
            keys = view[first].tolist()
            unseen = ~np.fromiter(map(seen.__contains__, keys), bool, len(keys))
            fresh = first[unseen]
            if not len(fresh):
                break
            if len(fresh) > max_configurations:
                truncated = True
                break
            seen.update(compress(keys, unseen.tolist()))
            rows, ns = rows[fresh], ns[fresh]
            t1 = time.perf_counter()
            _, _, _, canon = K.canonical_rows(rows, ns)
            canon[np.arange(ROW_WIDTH)[None, :] >= ns[:, None]] = PAD
            ckeys = _keys(canon)
            t2 = time.perf_counter()
            cset = set(ckeys)
            todo = sorted(cset.difference(verdicts))
            found = 0
            for lo in range(0, len(todo), SEARCH_BATCH_MAX):
                part = todo[lo:lo + SEARCH_BATCH_MAX]
                R = np.frombuffer(b"".join(part), np.uint8).reshape(len(part), ROW_WIDTH)
                N = (R != PAD).sum(axis=1).astype(np.int32)
                e0 = D.odd_cycle_counts(R, N)
                three = (N - e0) // 2
                res = D.sort_rows(slot, R, N, max_ratio, max_depth, e0=e0, norms=three)
                verdicts.update(zip(part, res.found.tolist()))
                norms.update(zip(part, three.tolist()))
                hit = np.flatnonzero(res.found).tolist()
                width = N.tolist()
                sortings.update((tuple(part[q][:width[q]]), res.moves[q]) for q in hit)
                found += len(hit)
            t3 = time.perf_counter()
            bad = np.flatnonzero(~np.fromiter(map(verdicts.__getitem__, ckeys), bool, len(ckeys))).tolist()
            last = depth + 1 == max_levels
            parts_r, parts_n = [], []
            tick = time.perf_counter()
            if not last and workers > 1 and len(bad) >= 2 * workers:
                if pool is None:
                    pool = Pool(workers)
                todo_rows = [rows[i, :ns[i]] for i in bad]
                step = max(1, len(todo_rows) // (8 * workers))
                chunks = [(extend_rows, todo_rows[lo:lo + step]) for lo in range(0, len(todo_rows), step)]
                for t, out in enumerate(pool.imap(_extend_chunk, chunks)):
                    for r_, n_ in out:
                        parts_r.append(r_)
                        parts_n.append(n_)
                    if log and time.perf_counter() - tick > 30:
                        tick = time.perf_counter()
                        log(f"level {depth}: extended {min(len(bad), (t + 1) * step)} of {len(bad)} configurations")
            elif not last:
                for t, i in enumerate(bad):
                    r_, n_ = extend_rows(rows[i, :ns[i]])
                    parts_r.append(r_)
                    parts_n.append(n_)
                    if log and time.perf_counter() - tick > 30:
                        tick = time.perf_counter()
                        log(f"level {depth}: extended {t + 1} of {len(bad)} configurations")
            t4 = time.perf_counter()
            lv = Level(depth, len(fresh), len(todo), found, len(bad), max(map(norms.__getitem__, cset)),
                       max((norms[ckeys[i]] for i in bad), default=-1),
                       {"dedup": t1 - t0, "canonical": t2 - t1, "search": t3 - t2, "extend": t4 - t3})
            levels.append(lv)
            if log:
                log(f"level {depth}: {lv.configurations} configurations, {lv.new_canonical} new canonical, "
                    f"{lv.found} found, {lv.extended} without a sorting (3-norm <= {lv.max_bad_three_norm}), "
                    f"3-norm <= {lv.max_three_norm}, "
                    + ", ".join(f"{k} {v:.2f}s" for k, v in lv.seconds.items()))
            if last or not parts_r:
                break
            rows, ns = np.concatenate(parts_r), np.concatenate(parts_n)
    finally:
        if pool is not None:
            pool.close()
            pool.join()
    pad = bytes([PAD])

    def symbols(k: bytes) -> Tuple[int, ...]:
        n = k.find(pad)
        return tuple(k if n < 0 else k[:n])

    return Exploration(levels, {symbols(k): v for k, v in verdicts.items()}, sortings, truncated)


def main(argv=None):
    """python -m tdp1375_amd.extensions --tree dfs --levels 3"""
    import argparse
    import json
    import sys

    from . import gpu_sorting_search as G
    ap = argparse.ArgumentParser(description="level-by-level SortOrExtend exploration on the GPU")
    ap.add_argument("--tree", choices=("dfs", "comb"), default="dfs")
    ap.add_argument("--levels", type=int, default=3)
    ap.add_argument("--max-ratio", type=float, default=1.375)
    ap.add_argument("--max-depth", type=int, default=11)
    ap.add_argument("--max-configurations", type=int, default=1 << 21)
    ap.add_argument("--queue-mb", type=int, default=16384)
    # 2^28 entries (19.3 GB of tags and keys): the exhaustive depth-11 rungs of
    # the unsatisfiable configurations re-expand fewer states than with 2^26
    ap.add_argument("--dedup-log2", type=int, default=28)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="host processes that extend a level's configurations")
    args = ap.parse_args(argv)
    roots = DFS_ROOTS if args.tree == "dfs" else [c for c, _ in BAD_SMALL_COMPONENTS]
    ext = dfs_extension_rows if args.tree == "dfs" else combination_extension_rows
    G.init_slots(1, 1, args.queue_mb << 20, 1 << args.dedup_log2)
    try:
        t0 = time.perf_counter()
        res = explore(roots, ext, args.levels, args.max_ratio, args.max_depth,
                      max_configurations=args.max_configurations,
                      log=lambda m: print(m, file=sys.stderr, flush=True), workers=args.workers)
        wall = time.perf_counter() - t0
    finally:
        G.destroy_slots()
    print(json.dumps({"tree": args.tree, "levels": [vars(l) for l in res.levels], "truncated": res.truncated,
                      "canonical_searched": len(res.verdicts),
                      "bad": sum(1 for v in res.verdicts.values() if not v), "seconds": wall}), flush=True)


if __name__ == "__main__":
    main()
