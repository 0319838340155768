"""Learned move ordering: data, features and training (host side).

The plain search orders the admitted children of a node by triple index
(sorting_search.cu:533-535).  The learned ordering scores each admitted child
with a small MLP (128 feature bits -> 256 ReLU -> 1 logit, bf16 weights, fp32
accumulate, MFMA on the GPU) and takes the best-scored one first; the set of
children, and so every verdict, is unchanged.

Training pairs come from the plain search itself: for every configuration it
found, each state s_t on the returned sequence (s_0 the start) is a group whose
rows are the children of s_t that the search admits (or accepts) at depth t+1,
labelled 1 for the move the sequence took.  The network is trained on the GPU
(tdp_mlp_train: forward/backward MFMA kernels and Adam) to maximise the
softmax probability of that move within its group.

The features of a move are computed on the device by the search kernel; this
module restates them with numpy for the training data (vectorised over all
C(n,3) triples of a state), and tests check both against each other.
"""
from __future__ import annotations

import ctypes
import functools
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

IN, HIDDEN = 128, 256
N_PARAMS = IN * HIDDEN + 2 * HIDDEN + 1

# bit offsets of the feature groups (tdp_learn.h move_features)
FEATURE_LAYOUT = {
    "delta": 0, "child_depth": 3, "gain": 19, "moves_left": 35, "len_a": 51, "len_b": 67, "len_c": 83,
    "sharing": 99, "arc_parity": 104, "cycle_order": 107, "n_quarter": 108, "adjacent_ij": 124,
    "adjacent_jk": 125, "i_first": 126, "k_last": 127,
}


@functools.lru_cache(maxsize=64)
def triples(n: int) -> np.ndarray:
    """All i<j<k in lexicographic order (the triple index order), shape (T, 3)."""
    if n < 3:
        return np.zeros((0, 3), np.int64)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m = (i < j) & (j < k)
    return np.stack([i[m], j[m], k[m]], axis=1)


def cycle_info(spi: Sequence[int]) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Per symbol: cycle id (smallest symbol), position from it along spi, length."""
    n = len(spi)
    root = np.zeros(n, np.int64)
    pos = np.zeros(n, np.int64)
    length = np.zeros(n, np.int64)
    seen = np.zeros(n, bool)
    for s in range(n):
        if seen[s]:
            continue
        cyc = [s]
        x = spi[s]
        while x != s:
            cyc.append(x)
            x = spi[x]
        r = min(cyc)
        t0 = cyc.index(r)
        cyc = cyc[t0:] + cyc[:t0]
        for p, x in enumerate(cyc):
            root[x], pos[x], length[x] = r, p, len(cyc)
            seen[x] = True
    return root, pos, length


def deltas_and_features(pi: Sequence[int], spi: Sequence[int], e0: int, d: int, D: int
                        ) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """For every triple (lexicographic order): the change of the odd-cycle
    count, the 4 feature words, and the triples.  Restates the kernel's
    parity_delta (tdp_common.h) and move_features (tdp_learn.h)."""
    n = len(pi)
    tr = triples(n)
    root, pos, length = cycle_info(spi)
    pi = np.asarray(pi, np.int64)
    a, b, c = pi[tr[:, 0]], pi[tr[:, 1]], pi[tr[:, 2]]
    ra, rb, rc = root[a], root[b], root[c]
    pa, pb, pc = pos[a], pos[b], pos[c]
    la, lb, lc = length[a], length[b], length[c]
    ab, bc, ac = ra == rb, rb == rc, ra == rc
    same = ab & bc
    kab, kbc, kca = (pb - pa) % la, (pc - pb) % la, (pa - pc) % la
    ordered = same & (kab + kbc + kca == la)
    d_same = np.where(ordered, (kab & 1) + (kbc & 1) + (kca & 1) - (la & 1), 0)
    d_three = ((la + lb + lc) & 1) - (la & 1) - (lb & 1) - (lc & 1)
    # two in one cycle (length L, positions x->y, k steps), third in another (Lo)
    L = np.where(ab, la, np.where(bc, lb, la))
    Lo = np.where(ab, lc, np.where(bc, la, lb))
    k = np.where(ab, (pb - pa) % np.maximum(la, 1), np.where(bc, (pc - pb) % np.maximum(lb, 1),
                                                              (pc - pa) % np.maximum(la, 1)))
    after = np.where(ac & ~ab & ~bc, ((L - k) & 1) + ((Lo + k) & 1), (k & 1) + ((L + Lo - k) & 1))
    d_two = after - (L & 1) - (Lo & 1)
    delta = np.where(same, d_same, np.where(~ab & ~bc & ~ac, d_three, d_two))

    e = int(np.sum([1 for r in set(root.tolist()) if length[r] % 2 == 1]))
    echild = e + delta
    bits = np.zeros((len(tr), IN), bool)
    rows = np.arange(len(tr))

    def one_hot(off, v):
        bits[rows, off + v] = True

    one_hot(0, (delta + 2) // 2)
    bits[:, 3 + min(max(d + 1, 0), 15)] = True
    one_hot(19, np.clip(echild - e0, 0, 15))
    bits[:, 35 + min(max(D - (d + 1), 0), 15)] = True
    one_hot(51, np.minimum(la, 16) - 1)
    one_hot(67, np.minimum(lb, 16) - 1)
    one_hot(83, np.minimum(lc, 16) - 1)
    pattern = np.where(same, 4, np.where(ab, 1, np.where(bc, 2, np.where(ac, 3, 0))))
    one_hot(99, pattern)
    bits[:, 104] = ab & (((pb - pa) % la) & 1 == 1)
    bits[:, 105] = bc & (((pc - pb) % lb) & 1 == 1)
    bits[:, 106] = ac & (((pa - pc) % lc) & 1 == 1)
    bits[:, 107] = ordered
    bits[:, 108 + min(n >> 2, 15)] = True
    bits[:, 124] = tr[:, 1] == tr[:, 0] + 1
    bits[:, 125] = tr[:, 2] == tr[:, 1] + 1
    bits[:, 126] = tr[:, 0] == 0
    bits[:, 127] = tr[:, 2] == n - 1
    return delta, pack_bits(bits), tr


def pack_bits(bits: np.ndarray) -> np.ndarray:
    """(rows, 128) bool -> (rows, 4) uint32, bit b in word b // 32."""
    w = np.zeros((bits.shape[0], 4), np.uint64)
    for q in range(4):
        chunk = bits[:, 32 * q:32 * q + 32].astype(np.uint64)
        w[:, q] = (chunk << np.arange(32, dtype=np.uint64)).sum(axis=1)
    return w.astype(np.uint32)


def predicate_tables(n: int, e0: int, min_rate: float, D: int, full: bool = False):
    """accept[d] / reach[d] bit tables of the search (tdp_host_predicate_tables)."""
    acc = (ctypes.c_uint64 * 16)()
    rch = (ctypes.c_uint64 * 16)()
    _lib.check(_lib.load().tdp_host_predicate_tables(n, e0, min_rate, D, int(full), acc, rch))
    return list(acc), list(rch)


def apply_positions(pi: List[int], spi: List[int], i: int, j: int, k: int) -> None:
    """One move on (pi, spi) in place: spi' = spi o (a b c)^-1 with
    a = pi[i], b = pi[j], c = pi[k] (sorting_search.cu:547-557), and the block
    swap of pi[i:j] and pi[j:k] (transposition.cuh:13-19)."""
    a, b, c = pi[i], pi[j], pi[k]
    sa, sb, sc = spi[a], spi[b], spi[c]
    spi[a], spi[b], spi[c] = sc, sa, sb
    pi[i:k] = pi[j:k] + pi[i:j]


def odd_cycles(spi: Sequence[int]) -> int:
    root, _, length = cycle_info(spi)
    return int(sum(1 for s in range(len(spi)) if root[s] == s and length[s] % 2 == 1))


@dataclass
class Dataset:
    rows: np.ndarray          # (N, 4) uint32 feature words
    labels: np.ndarray        # (N,) uint8
    group_start: np.ndarray   # (G + 1,) int32

    @property
    def groups(self) -> int:
        return len(self.group_start) - 1


def build_dataset(cases, min_rate: float, max_depth: int, witnesses, max_rows_per_group: int = 4096) -> Dataset:
    """Groups of candidate moves along every found sequence (positions per
    move).  A group holds the children admitted or accepted at depth t+1 by
    the search's predicates; groups whose taken move is not among them, or
    that exceed max_rows_per_group, are skipped."""
    rows, labels, starts = [], [], [0]
    for case, moves in zip(cases, witnesses):
        if not moves:
            continue
        pi, spi = list(case.pi), list(case.spi)
        n = len(pi)
        e0 = case.even_cycles
        acc, rch = predicate_tables(n, e0, min_rate, max_depth)
        for t, (i, j, k) in enumerate(moves):
            e = odd_cycles(spi)
            delta, feats, tr = deltas_and_features(pi, spi, e0, t, max_depth)
            ec = e + delta
            ok = ((np.uint64(acc[t + 1]) >> ec.astype(np.uint64)) & np.uint64(1)).astype(bool)
            if t + 1 < max_depth:
                ok |= ((np.uint64(rch[t + 1]) >> ec.astype(np.uint64)) & np.uint64(1)).astype(bool)
            lab = (tr[:, 0] == i) & (tr[:, 1] == j) & (tr[:, 2] == k)
            if lab[ok].sum() == 1 and 1 < ok.sum() <= max_rows_per_group:
                rows.append(feats[ok])
                labels.append(lab[ok].astype(np.uint8))
                starts.append(starts[-1] + int(ok.sum()))
            apply_positions(pi, spi, i, j, k)
    if not rows:
        return Dataset(np.zeros((0, 4), np.uint32), np.zeros(0, np.uint8), np.zeros(1, np.int32))
    return Dataset(np.concatenate(rows), np.concatenate(labels), np.asarray(starts, np.int32))


# ---------------------------------------------------------------- model (C ABI)
def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def init(seed: int = 1) -> None:
    _lib.check(_lib.load().tdp_mlp_init(seed))


def set_weights(w1: np.ndarray, b1: np.ndarray, w2: np.ndarray, b2: float) -> None:
    w1, b1, w2 = _f32(w1).reshape(IN, HIDDEN), _f32(b1), _f32(w2)
    _lib.check(_lib.load().tdp_mlp_set_weights(_ptr(w1, ctypes.c_float), _ptr(b1, ctypes.c_float),
                                               _ptr(w2, ctypes.c_float), float(b2)))


def get_weights():
    w1 = np.zeros((IN, HIDDEN), np.float32)
    b1 = np.zeros(HIDDEN, np.float32)
    w2 = np.zeros(HIDDEN, np.float32)
    b2 = ctypes.c_float()
    _lib.check(_lib.load().tdp_mlp_get_weights(_ptr(w1, ctypes.c_float), _ptr(b1, ctypes.c_float),
                                               _ptr(w2, ctypes.c_float), ctypes.byref(b2)))
    return w1, b1, w2, float(b2.value)


def forward(rows: np.ndarray) -> np.ndarray:
    rows = np.ascontiguousarray(rows, np.uint32)
    out = np.zeros(len(rows), np.float32)
    if len(rows):
        _lib.check(_lib.load().tdp_mlp_forward(_ptr(rows, ctypes.c_uint32), len(rows), _ptr(out, ctypes.c_float)))
    return out


def gradients(ds: Dataset):
    """(flat gradient [w1 | b1 | w2 | b2], mean loss) of one full batch."""
    g = np.zeros(N_PARAMS, np.float32)
    loss = ctypes.c_float()
    rows = np.ascontiguousarray(ds.rows, np.uint32)
    lab = np.ascontiguousarray(ds.labels, np.uint8)
    gs = np.ascontiguousarray(ds.group_start, np.int32)
    _lib.check(_lib.load().tdp_mlp_gradients(_ptr(rows, ctypes.c_uint32), _ptr(lab, ctypes.c_uint8),
                                             _ptr(gs, ctypes.c_int32), ds.groups, _ptr(g, ctypes.c_float),
                                             ctypes.byref(loss)))
    return g, float(loss.value)


def train(ds: Dataset, epochs: int = 4, lr: float = 1e-3, groups_per_batch: int = 256, seed: int = 7
          ) -> np.ndarray:
    """Adam on the GPU; returns the mean loss of every epoch."""
    losses = np.zeros(epochs, np.float32)
    rows = np.ascontiguousarray(ds.rows, np.uint32)
    lab = np.ascontiguousarray(ds.labels, np.uint8)
    gs = np.ascontiguousarray(ds.group_start, np.int32)
    _lib.check(_lib.load().tdp_mlp_train(_ptr(rows, ctypes.c_uint32), _ptr(lab, ctypes.c_uint8),
                                         _ptr(gs, ctypes.c_int32), ds.groups, epochs, lr, groups_per_batch, seed,
                                         _ptr(losses, ctypes.c_float)))
    return losses


def group_accuracy(ds: Dataset, logits: np.ndarray) -> float:
    """Share of groups whose labelled row has the largest logit."""
    hit = 0
    for g in range(ds.groups):
        s, e = ds.group_start[g], ds.group_start[g + 1]
        hit += int(np.argmax(logits[s:e]) == np.argmax(ds.labels[s:e]))
    return hit / max(ds.groups, 1)
