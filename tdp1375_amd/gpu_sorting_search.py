"""Host-side mirror of br.unb.cic.tdp.base.GPUSortingSearch over the C ABI.

Same constructor arguments, method names, argument meaning and error
behaviour as the reference facade (GPUSortingSearch.java:13-87):

* the constructor creates ``devices_count * slots_per_device`` slots and a
  blocking slot pool (ArrayBlockingQueue in the reference);
* ``search`` takes a slot, runs the native search, returns the moves as symbol
  triples (a, b, c) obtained by replaying the returned positions on a copy of
  ``pi`` (GPUSortingSearch.java:65-75), and returns the slot;
* native failures raise RuntimeError subclasses carrying the native message
  ("GPU sorting search cancelled", "GPU sorting search queue overflow", ...).
"""
from __future__ import annotations

import ctypes
import queue
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import CancelledError, QueueOverflowError, TdpError  # noqa: F401  (re-export)


def apply_transposition(pi: bytearray, i: int, j: int, k: int) -> None:
    """Block swap of pi[i:j] and pi[j:k] (GPUSortingSearch.java:81-87)."""
    pi[i:k] = pi[j:k] + pi[i:j]


def positions_to_symbols(pi: Sequence[int], positions) -> List[Tuple[int, int, int]]:
    cur = bytearray(int(x) & 0xFF for x in pi)
    moves = []
    for i, j, k in positions:
        moves.append((cur[i], cur[j], cur[k]))
        apply_transposition(cur, i, j, k)
    return moves


def init_slots(total_slots: int, devices_count: int, queue_bytes_budget: int, dedup_table_size: int) -> None:
    _lib.check(_lib.load().tdp_init_slots(total_slots, devices_count, queue_bytes_budget, dedup_table_size))


def destroy_slots() -> None:
    _lib.check(_lib.load().tdp_destroy_slots())


def cancel_sort(slot: int) -> None:
    _lib.load().tdp_cancel(slot)


@dataclass
class NativeSearchResult:
    positions: List[Tuple[int, int, int]]
    counters: dict


def search_for_sorting_seq(pi: Sequence[int], spi: Sequence[int], initial_even_cycles: int, min_rate: float,
                           max_moves: int, full_sorting: bool, slot: int, flags: int = 0) -> NativeSearchResult:
    """The native call (JNI searchForSortingSeq): positions (i, j, k) per move."""
    n = len(pi)
    if len(spi) < n:
        raise TdpError(_lib.TDP_ERR_INVALID_ARGUMENT, "performGpuSortingSearch: inconsistent input sizes")
    L = _lib.load()
    moves = (ctypes.c_int32 * (3 * _lib.MAX_DEPTH))()
    depth = ctypes.c_int32(0)
    cnt = _lib.Counters()
    st = L.tdp_search_ex(slot, _lib.u8(pi, 64), _lib.u8(spi[:n], 64), n, int(initial_even_cycles),
                         float(min_rate), int(max_moves), int(bool(full_sorting)), flags, moves,
                         ctypes.byref(depth), ctypes.byref(cnt))
    _lib.check(st)
    pos = [(moves[3 * m], moves[3 * m + 1], moves[3 * m + 2]) for m in range(depth.value)]
    return NativeSearchResult(pos, cnt.as_dict())


def filter_pi_permutations(cycle: Sequence[int], cycle_len: int, oriented_triple: Sequence[int], triple_len: int,
                           n: int, skip_2move_check: bool, max_rows: int = _lib.FILTER_MAX_OUTPUT):
    """GPUSortingSearch.filterPiPermutations: the first (at most 2^20) passing
    arrangements in rank order, as a (rows, n) uint8 array, plus the uncapped count."""
    L = _lib.load()
    if cycle_len < 0 or triple_len < 0 or len(cycle) < cycle_len or len(oriented_triple) < triple_len:
        raise TdpError(_lib.TDP_ERR_INVALID_ARGUMENT, "invalid cycle or triple length")
    rows = max(0, int(max_rows)) if 1 <= n <= 63 else 0
    out = np.zeros((max(rows, 1), max(n, 1)), np.uint8)
    total = ctypes.c_int64(0)
    written = ctypes.c_int64(0)
    st = L.tdp_filter(_lib.u8(cycle[:cycle_len]), cycle_len, _lib.u8(oriented_triple[:triple_len]), triple_len, n,
                      int(bool(skip_2move_check)), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), rows,
                      ctypes.byref(total), ctypes.byref(written))
    _lib.check(st)
    return out[: written.value].copy(), int(total.value)


def filter_last_kernel_ms() -> float:
    return float(_lib.load().tdp_filter_last_kernel_ms())


def set_tuning(waves_per_cu: int = 0, share_interval: int = 0, share_max: int = 0, idle_backoff: int = 0,
               max_active: int = 0) -> None:
    """Work-sharing launch tuning for later launches (0 keeps a default)."""
    _lib.check(_lib.load().tdp_set_tuning(waves_per_cu, share_interval, share_max, idle_backoff, max_active))


def set_expansion_batch(nodes: int = 0) -> None:
    """Nodes a wave expands before one dedup pass (work-sharing mode; 0 =
    default 12, at most 16)."""
    _lib.check(_lib.load().tdp_set_expansion_batch(int(nodes)))


def set_launch_deadline(milliseconds: int) -> None:
    """Time budget of each launch run with FLAG_SOFT_DEADLINE."""
    _lib.check(_lib.load().tdp_set_launch_deadline(int(milliseconds)))


def last_launch_stats(slot: int) -> dict:
    s = _lib.LaunchStats()
    _lib.check(_lib.load().tdp_last_launch_stats(slot, ctypes.byref(s)))
    out = {name: getattr(s, name) for name, _ in s._fields_ if name != "phase_ticks"}
    out["phase_ms"] = {k: s.phase_ticks[q] * 1e-5 for q, k in enumerate(_lib.PHASES)}
    return out


@dataclass
class BatchResult:
    found: np.ndarray          # (count,) bool
    depth: np.ndarray          # (count,) int
    moves: list                # per case: list of (i, j, k) positions (None: see moves_array)
    counters: dict             # name -> (count,) uint64 array
    moves_array: Optional[np.ndarray] = None   # (count, MAX_DEPTH, 3) positions, rows beyond depth unset


def search_batch(slot: int, pis: Sequence[Sequence[int]], spis: Sequence[Sequence[int]],
                 initial_even_cycles: Sequence[int], min_rates, max_moves, full_sorting=None,
                 flags: int = 0) -> BatchResult:
    """Many configurations in one persistent launch (tdp_search_batch)."""
    count = len(pis)
    P = np.zeros((count, 64), np.uint8)
    S = np.zeros((count, 64), np.uint8)
    ns = np.zeros(count, np.int32)
    for c in range(count):
        n = len(pis[c])
        ns[c] = n
        P[c, :n] = np.asarray(pis[c], np.uint8)
        S[c, :n] = np.asarray(spis[c][:n], np.uint8)
    return search_batch_rows(slot, P, S, ns, initial_even_cycles, min_rates, max_moves, full_sorting, flags)


_CASE_DTYPE = np.dtype(_lib.CaseResult)


def search_batch_rows(slot: int, P: np.ndarray, S: np.ndarray, ns, initial_even_cycles, min_rates, max_moves,
                      full_sorting=None, flags: int = 0, move_lists: bool = True) -> BatchResult:
    """search_batch on pre-packed rows: P and S (count, 64) uint8 (pi and spi
    of configuration c in their first ns[c] entries); moves are position
    triples, read with the found flags, depths and counters through a numpy
    view of the result array.  move_lists=False leaves moves None and gives
    the positions only as moves_array (no per-configuration Python lists)."""
    count = len(ns)
    P = np.ascontiguousarray(P, np.uint8)
    S = np.ascontiguousarray(S, np.uint8)
    nsa = np.ascontiguousarray(ns, np.int32)
    e0 = np.ascontiguousarray(initial_even_cycles, np.int32)
    mr = np.ascontiguousarray(np.broadcast_to(np.float32(min_rates) if np.isscalar(min_rates) else
                                              np.asarray(min_rates, np.float32), (count,)), np.float32)
    md = np.ascontiguousarray(np.broadcast_to(np.asarray(max_moves, np.int32), (count,)), np.int32)
    fs = np.ascontiguousarray(np.broadcast_to(np.asarray(0 if full_sorting is None else full_sorting, np.int32),
                                              (count,)), np.int32)
    out = (_lib.CaseResult * max(count, 1))()
    u8p = ctypes.POINTER(ctypes.c_uint8)
    i32p = ctypes.POINTER(ctypes.c_int32)
    st = _lib.load().tdp_search_batch(slot, count, P.ctypes.data_as(u8p), S.ctypes.data_as(u8p),
                                      nsa.ctypes.data_as(i32p), e0.ctypes.data_as(i32p),
                                      mr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), md.ctypes.data_as(i32p),
                                      fs.ctypes.data_as(i32p), flags, out)
    _lib.check(st)
    arr = np.frombuffer(out, dtype=_CASE_DTYPE, count=count)
    found = arr["found"] != 0
    depth = arr["depth"].astype(np.int32)
    mv = np.array(arr["moves"]).reshape(count, _lib.MAX_DEPTH, 3)
    moves = [[tuple(m) for m in mv[c, :max(int(depth[c]), 0)].tolist()] for c in range(count)] if move_lists else None
    counters = {name: arr["counters"][name].astype(np.uint64) for name, _ in _lib.Counters._fields_}
    return BatchResult(found, depth, moves, counters, mv)


class GPUSortingSearch:
    """Mirror of the Java facade (GPUSortingSearch.java:13-87)."""

    def __init__(self, devices_count: int, slots_per_device: int, queue_bytes_budget: int, dedup_table_size: int):
        total = devices_count * slots_per_device
        self._slot_pool: "queue.Queue[int]" = queue.Queue(maxsize=max(total, 1))
        for i in range(total):
            self._slot_pool.put(i)
        init_slots(total, devices_count, queue_bytes_budget, dedup_table_size)
        self.last_counters: Optional[dict] = None

    destroy_slots = staticmethod(destroy_slots)
    cancel_sort = staticmethod(cancel_sort)
    filter_pi_permutations = staticmethod(filter_pi_permutations)

    def search(self, pi: Sequence[int], spi: Sequence[int], initial_even_cycles: int, min_rate: float,
               max_moves: int, full_sorting: bool, flags: int = 0) -> List[Tuple[int, int, int]]:
        slot = self._slot_pool.get()
        try:
            res = search_for_sorting_seq(pi, spi, initial_even_cycles, min_rate, max_moves, full_sorting, slot,
                                         flags)
            self.last_counters = res.counters
            if not res.positions:
                return []
            return positions_to_symbols(pi, res.positions)
        finally:
            self._slot_pool.put(slot)
