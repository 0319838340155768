"""ctypes wrapper around the CPU oracle (oracle/tdp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the CPU baseline.  The product
package (tdp1375_amd) never imports it.

Each function restates the reference algorithm; see the header of
tdp_oracle.c for the file:line each one follows and the checks that pin it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libtdp_oracle.so")
MAX_DEPTH = 15


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


class _Result(ctypes.Structure):
    _fields_ = [
        ("found", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("moves", ctypes.c_int32 * (MAX_DEPTH * 3)),
        ("aborted", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("expansions", ctypes.c_uint64),
        ("generated", ctypes.c_uint64),
        ("admitted", ctypes.c_uint64),
        ("dominated", ctypes.c_uint64),
        ("improved", ctypes.c_uint64),
        ("inserted", ctypes.c_uint64),
        ("roots_live", ctypes.c_uint64),
        ("pops", ctypes.c_uint64),
    ]


COUNTER_NAMES = ("expansions", "generated", "admitted", "dominated", "improved",
                 "inserted", "roots_live")


@dataclass
class SearchResult:
    found: bool
    depth: int
    moves: List[Tuple[int, int, int]]
    aborted: bool
    counters: dict


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.or_count_odd_cycles.argtypes = [u8p, ctypes.c_int]
        L.or_satisfies_rate.argtypes = [ctypes.c_int] * 4 + [ctypes.c_float, ctypes.c_int]
        L.or_can_still_reach.argtypes = [ctypes.c_int] * 4 + [ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.or_triple_count.argtypes = [ctypes.c_int]
        L.or_triple_unrank.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p]
        L.or_apply_move.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_search.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_Result)]
        L.or_search_many.argtypes = [ctypes.c_int, u8p, u8p, i32p, i32p, ctypes.POINTER(ctypes.c_float),
                                     i32p, i32p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.POINTER(_Result)]
        L.or_search_many.restype = ctypes.c_double
        L.or_search_ex.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                   ctypes.POINTER(_Result)]
        L.or_search_many_ex.argtypes = [ctypes.c_int, u8p, u8p, i32p, i32p, ctypes.POINTER(ctypes.c_float),
                                        i32p, i32p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(_Result)]
        L.or_search_many_ex.restype = ctypes.c_double
        L.or_children_even_arcs.argtypes = [u8p, u8p, ctypes.c_int, i32p]
        L.or_java_full_sort_search.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                               ctypes.c_int, ctypes.c_int, i32p]
        L.or_bfs_verdict.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                     ctypes.c_int]
        L.or_replay_positions.argtypes = [u8p, u8p, ctypes.c_int, i32p, ctypes.c_int]
        L.or_replay_symbols.argtypes = [u8p, u8p, ctypes.c_int, i32p, ctypes.c_int]
        L.or_filter.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, u8p,
                                ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.or_filter.restype = ctypes.c_int64
        L.or_children_even.argtypes = [u8p, u8p, ctypes.c_int, i32p]
        L.or_filter_accepts.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _u8(seq: Sequence[int], size: Optional[int] = None):
    arr = (ctypes.c_uint8 * (size or max(1, len(seq))))()
    for i, v in enumerate(seq):
        arr[i] = int(v) & 0xFF
    return arr


def count_odd_cycles(spi: Sequence[int]) -> int:
    return lib().or_count_odd_cycles(_u8(spi, 64), len(spi))


def satisfies_rate(e: int, d: int, n: int, e0: int, min_rate: float, full: bool) -> bool:
    return bool(lib().or_satisfies_rate(e, d, n, e0, min_rate, int(full)))


def can_still_reach(e: int, d: int, n: int, e0: int, min_rate: float, max_depth: int, full: bool) -> bool:
    return bool(lib().or_can_still_reach(e, d, n, e0, min_rate, max_depth, int(full)))


def triple_count(n: int) -> int:
    return lib().or_triple_count(n)


def apply_move(pi: List[int], spi: List[int], i: int, j: int, k: int):
    n = len(pi)
    p, s = _u8(pi, 64), _u8(spi, 64)
    lib().or_apply_move(p, s, n, i, j, k)
    return list(p[:n]), list(s[:n])


def children_even(pi, spi, arcs: bool = False) -> np.ndarray:
    """Odd-cycle count of every child, lexicographic triple order: by an O(n)
    recount of each child, or (arcs=True) by re-linking the arcs of the three
    touched cycles."""
    n = len(pi)
    T = max(1, triple_count(n))
    out = np.zeros(T, np.int32)
    fn = lib().or_children_even_arcs if arcs else lib().or_children_even
    fn(_u8(pi, 64), _u8(spi, 64), n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out[: triple_count(n)]


def _unpack(r: _Result) -> SearchResult:
    moves = [(r.moves[3 * m], r.moves[3 * m + 1], r.moves[3 * m + 2]) for m in range(r.depth)] if r.found else []
    return SearchResult(bool(r.found), int(r.depth) if r.found else 0, moves, bool(r.aborted),
                        {k: int(getattr(r, k)) for k in COUNTER_NAMES})


ARC_CHILDREN = 1   # OR_FLAG_ARC_CHILDREN


def search(pi, spi, e0: int, min_rate: float, max_depth: int, full: bool = False,
           use_dedup: bool = True, max_expansions: int = 0, arcs: bool = False) -> SearchResult:
    """Deterministic single-queue search.  arcs=True evaluates children by arc
    re-linking instead of the O(n) recount (same results, faster)."""
    r = _Result()
    st = lib().or_search_ex(_u8(pi, 64), _u8(spi, 64), len(pi), e0, min_rate, max_depth, int(full),
                            int(use_dedup), max_expansions, ARC_CHILDREN if arcs else 0, ctypes.byref(r))
    if st != 0:
        raise ValueError(f"oracle search rejected arguments ({st})")
    return _unpack(r)


def search_many(pis: np.ndarray, spis: np.ndarray, ns, e0s, min_rates, max_depths, fulls,
                use_dedup: bool = True, max_expansions: int = 0, threads: int = 1, arcs: bool = False):
    """Runs independent searches on `threads` host threads; returns (results, seconds).
    arcs=True: children evaluated by arc re-linking (see search)."""
    count = len(ns)
    P = np.zeros((count, 64), np.uint8)
    S = np.zeros((count, 64), np.uint8)
    for c in range(count):
        P[c, :ns[c]] = pis[c][:ns[c]]
        S[c, :ns[c]] = spis[c][:ns[c]]
    ns_ = np.ascontiguousarray(ns, np.int32)
    e0_ = np.ascontiguousarray(e0s, np.int32)
    mr_ = np.ascontiguousarray(min_rates, np.float32)
    md_ = np.ascontiguousarray(max_depths, np.int32)
    fu_ = np.ascontiguousarray(fulls, np.int32)
    outs = (_Result * count)()
    u8p = ctypes.POINTER(ctypes.c_uint8)
    i32p = ctypes.POINTER(ctypes.c_int32)
    secs = lib().or_search_many_ex(count, P.ctypes.data_as(u8p), S.ctypes.data_as(u8p),
                                   ns_.ctypes.data_as(i32p), e0_.ctypes.data_as(i32p),
                                   mr_.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   md_.ctypes.data_as(i32p), fu_.ctypes.data_as(i32p), int(use_dedup),
                                   max_expansions, ARC_CHILDREN if arcs else 0, threads, outs)
    return [_unpack(outs[c]) for c in range(count)], secs


def java_full_sort_search(pi, spi, e0: int, min_rate: float, max_moves: int, float_prune: bool = False):
    out = (ctypes.c_int32 * (3 * MAX_DEPTH))()
    d = lib().or_java_full_sort_search(_u8(pi, 64), _u8(spi, 64), len(pi), e0, min_rate, max_moves,
                                       int(float_prune), out)
    if d < 0:
        raise ValueError("invalid arguments")
    return [(out[3 * m], out[3 * m + 1], out[3 * m + 2]) for m in range(d)]


def bfs_verdict(pi, spi, e0: int, min_rate: float, max_depth: int, full: bool = False) -> bool:
    v = lib().or_bfs_verdict(_u8(pi, 64), _u8(spi, 64), len(pi), e0, min_rate, max_depth, int(full))
    if v < 0:
        raise ValueError("bfs oracle supports n <= 9")
    return bool(v)


def replay_positions(pi, spi, moves):
    n = len(pi)
    p, s = _u8(pi, 64), _u8(spi, 64)
    flat = (ctypes.c_int32 * max(3, 3 * len(moves)))(*[int(x) for m in moves for x in m])
    e = lib().or_replay_positions(p, s, n, flat, len(moves))
    return e, list(p[:n]), list(s[:n])


def replay_symbols(pi, spi, moves):
    n = len(pi)
    p, s = _u8(pi, 64), _u8(spi, 64)
    flat = (ctypes.c_int32 * max(3, 3 * len(moves)))(*[int(x) for m in moves for x in m])
    e = lib().or_replay_symbols(p, s, n, flat, len(moves))
    return e, list(p[:n]), list(s[:n])


def filter_pi(cycle, triple, n: int, skip_2move_check: bool = False, out_cap: int = 1 << 20,
              threads: int = 1):
    """Returns (uncapped count, first out_cap passing arrangements in rank order)."""
    out = np.zeros((max(1, out_cap), n), np.uint8)
    written = ctypes.c_int64(0)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    cnt = lib().or_filter(_u8(cycle), len(cycle), _u8(triple), len(triple), n, int(skip_2move_check),
                          out.ctypes.data_as(u8p), out_cap, threads, ctypes.byref(written))
    if cnt < 0:
        raise ValueError("invalid filter arguments")
    return int(cnt), out[: written.value].copy()


def filter_accepts(pi, cycle, triple, skip_2move_check: bool = False) -> bool:
    return bool(lib().or_filter_accepts(_u8(pi, 64), len(pi), _u8(cycle), len(cycle), _u8(triple),
                                        len(triple), int(skip_2move_check)))
