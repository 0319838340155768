"""Proof-generation driver over the GPU search.

Host-side restatement of the caller of the search path:

* ``SortOrExtend.searchForSorting`` (SortOrExtend.java:89-111): iterative
  deepening at D = 4, then min(ub, 8) when ub > 4, then min(ub, max_depth)
  when ub > 8, with ub = floor(3-norm(spi) * (2 / minRate)) and the identity
  pi;
* ``SortOrExtend.compute`` (SortOrExtend.java:53-87): single flight per
  configuration key through ``SortingCoordinator``, resume from the files of
  an earlier run (a sorting file means done, a ``bad-cases/<spi>`` marker
  means searched without success), and ``extend`` only when this call searched
  and found nothing;
* ``SortingCoordinator.tryCompute`` (SortingCoordinator.java:72-98).

``sort_batch`` is the MI355X form of the same ladder: every configuration of a
list runs at its first depth in one batched launch, the ones still open run at
their second depth in the next, and so on, so a whole extension level costs
three launches instead of three per configuration.

The key is the spi written as the reference writes it
(``MulticyclePermutation.toString``: cycles in the given order, each from its
smallest symbol).  Canonicalisation of configurations (``Configuration
.getCanonical``) is not part of this package: callers pass configurations in
the form they want deduplicated.  The sorting file is plain text (one move
``(a b c)`` per line) instead of the reference's Velocity-rendered HTML; an
existing ``<spi>.html`` of a reference run is honoured on resume.
"""
from __future__ import annotations

import os
import threading
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .gpu_sorting_search import GPUSortingSearch, search_batch_rows

Cycles = Sequence[Sequence[int]]
Move = Tuple[int, int, int]


# ---------------------------------------------------------------- permutations
def spi_string(cycles: Cycles) -> str:
    """MulticyclePermutation.toString (MulticyclePermutation.java:111-116,
    Cycle.java:100-114): each cycle from its smallest symbol, in list order."""
    parts = []
    for c in cycles:
        c = list(c)
        m = c.index(min(c))
        parts.append("(" + " ".join(str(x) for x in c[m:] + c[:m]) + ")")
    return "".join(parts) if parts else "()"


def one_line(cycles: Cycles) -> List[int]:
    """MulticyclePermutation.oneLineNotation (MulticyclePermutation.java:60-72):
    the spi of n = total cycle length symbols, identity where no cycle maps."""
    n = sum(len(c) for c in cycles)
    out = list(range(n))
    for c in cycles:
        for t, s in enumerate(c):
            out[s] = c[(t + 1) % len(c)]
    return out


def odd_cycles(cycles: Cycles) -> int:
    """getNumberOfEvenCycles (MulticyclePermutation.java:149-151): the
    reference calls a cycle even when its length is odd (Cycle.java:148-150)."""
    return sum(1 for c in cycles if len(c) % 2 == 1)


def three_norm(cycles: Cycles) -> int:
    """MulticyclePermutation.get3Norm (MulticyclePermutation.java:161-163)."""
    return (sum(len(c) for c in cycles) - odd_cycles(cycles)) // 2


def upper_bound(cycles: Cycles, min_rate: float) -> float:
    """floor(3-norm * (2 / minRate)) with the float32 quotient of
    SortOrExtend.java:91."""
    q = np.float32(2.0) / np.float32(min_rate)
    return float(np.floor(np.float32(three_norm(cycles)) * q))


def depth_ladder(ub: float, max_depth: int) -> List[int]:
    """Depths tried by SortOrExtend.searchForSorting (SortOrExtend.java:95-105)."""
    ladder = [4]
    if ub > 4:
        ladder.append(int(min(ub, 8)))
    if ub > 8:
        ladder.append(int(min(ub, max_depth)))
    return ladder


# ---------------------------------------------------------------- single flight
class ComputeResult:
    COMPUTED, ALREADY_IN_PROGRESS, NOT_FOUND = "computed", "already_in_progress", "not_found"

    def __init__(self, kind: str, value=None):
        self.kind = kind
        self._value = value

    @property
    def is_computed(self) -> bool:
        return self.kind == self.COMPUTED

    @property
    def is_already_in_progress(self) -> bool:
        return self.kind == self.ALREADY_IN_PROGRESS

    @property
    def is_not_found(self) -> bool:
        return self.kind == self.NOT_FOUND

    @property
    def value(self):
        if self.kind != self.COMPUTED:
            raise RuntimeError(f"No value available for {self.kind}")
        return self._value


class SortingCoordinator:
    """Single flight per key (SortingCoordinator.java:72-98): the first caller
    computes, a concurrent caller with the same key returns
    ALREADY_IN_PROGRESS at once; the key is released when the computation
    ends, so a later call computes again."""

    def __init__(self):
        self._lock = threading.Lock()
        self._in_flight: Dict[str, Future] = {}

    def try_compute(self, key: str, compute: Callable[[], Optional[object]]) -> ComputeResult:
        fut: Future = Future()
        with self._lock:
            if key in self._in_flight:
                return ComputeResult(ComputeResult.ALREADY_IN_PROGRESS)
            self._in_flight[key] = fut
        try:
            result = compute()
            fut.set_result(result)
            return (ComputeResult(ComputeResult.COMPUTED, result) if result is not None
                    else ComputeResult(ComputeResult.NOT_FOUND))
        except BaseException as exc:
            fut.set_exception(exc)
            raise
        finally:
            with self._lock:
                if self._in_flight.get(key) is fut:
                    del self._in_flight[key]


# ---------------------------------------------------------------- driver
class SortOrExtend:
    """SortOrExtend over a GPUSortingSearch (SortOrExtend.java:27-111)."""

    coordinator = SortingCoordinator()

    def __init__(self, gpu: GPUSortingSearch, max_ratio: float, max_depth: int, output_dir: str,
                 extend: Optional[Callable[[Cycles], None]] = None):
        self.gpu = gpu
        self.min_rate = float(np.float32(2.0) / np.float32(max_ratio))   # SortOrExtend.java:44
        self.max_depth = max_depth
        self.output_dir = output_dir
        self.extend = extend
        self.searches = 0
        os.makedirs(os.path.join(output_dir, "bad-cases"), exist_ok=True)

    def sorting_path(self, key: str) -> str:
        return os.path.join(self.output_dir, key + ".txt")

    def _done(self, key: str) -> bool:
        return (os.path.exists(self.sorting_path(key))
                or os.path.exists(os.path.join(self.output_dir, key + ".html")))

    def compute(self, cycles: Cycles) -> ComputeResult:
        key = spi_string(cycles)

        def work():
            if self._done(key):
                return True
            bad = os.path.join(self.output_dir, "bad-cases", key)
            if not os.path.exists(bad):
                sorting = self.search_for_sorting(cycles)
                if sorting is not None:
                    write_sorting(self.sorting_path(key), sorting)
                    return True
                open(bad, "w").close()
            return None

        result = self.coordinator.try_compute(key, work)
        if result.is_not_found and self.extend is not None:
            self.extend(cycles)
        return result

    def search_for_sorting(self, cycles: Cycles) -> Optional[List[Move]]:
        """SortOrExtend.java:89-111: symbol moves of the first depth of the
        ladder that has a sequence, or None."""
        spi = one_line(cycles)
        pi = list(range(len(spi)))
        e0 = odd_cycles(cycles)
        for depth in depth_ladder(upper_bound(cycles, self.min_rate), self.max_depth):
            self.searches += 1
            moves = self.gpu.search(pi, spi, e0, self.min_rate, depth, False)
            if moves:
                return [tuple(int(x) for x in m) for m in moves]
        return None


def write_sorting(path: str, moves: Iterable[Move]) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        for a, b, c in moves:
            fh.write(f"({a} {b} {c})\n")
    os.replace(tmp, path)


@dataclass
class LadderResult:
    found: np.ndarray            # (count,) bool
    depth_tried: np.ndarray      # (count,) last depth searched
    moves: List[Optional[List[Move]]]   # symbol moves per configuration
    launches: int


def sort_batch(slot: int, configs: Sequence[Cycles], max_ratio: float, max_depth: int,
               flags: int = 0) -> LadderResult:
    """The ladder of SortOrExtend.searchForSorting for a whole list of
    configurations, one batched launch per rung."""
    count = len(configs)
    rows = np.zeros((count, 64), np.uint8)
    ns = np.zeros(count, np.int32)
    for c, cfg in enumerate(configs):
        spi = one_line(cfg)
        ns[c] = len(spi)
        rows[c, :len(spi)] = spi
    return sort_rows(slot, rows, ns, max_ratio, max_depth, e0=[odd_cycles(c) for c in configs],
                     norms=[three_norm(c) for c in configs], flags=flags)


def odd_cycle_counts(rows: np.ndarray, ns) -> np.ndarray:
    """Odd-length cycles (fixed points included) of every one-line row's
    first ns[c] entries, all rows at once: each symbol's cycle minimum by
    pointer doubling, cycle lengths by counting symbols per minimum."""
    rows = np.asarray(rows)
    ns = np.asarray(ns)
    K = rows.shape[0]
    if K == 0:
        return np.zeros(0, np.int64)
    # columns past the longest row are fixed points of every row: drop them
    W = max(1, int(ns.max()))
    rows = rows[:, :W]
    x = np.arange(W)
    inside = x[None, :] < ns[:, None]
    q = np.where(inside, rows.astype(np.intp), x[None, :])
    m = np.broadcast_to(x, (K, W)).copy()
    for _ in range(max(1, (W - 1).bit_length())):
        m = np.minimum(m, np.take_along_axis(m, q, 1))
        q = np.take_along_axis(q, q, 1)
    length = np.bincount((np.arange(K)[:, None] * W + m).ravel(), minlength=K * W).reshape(K, W)
    return ((length % 2 == 1) & (m == x[None, :]) & inside).sum(axis=1)


def _block_swap_rows(cur: np.ndarray, i, j, k) -> np.ndarray:
    """pi[:i] + pi[j:k] + pi[i:j] + pi[k:] for every row at once."""
    x = np.arange(cur.shape[1])[None, :]
    i, j, k = (np.asarray(v)[:, None] for v in (i, j, k))
    head = k - j
    src = np.where((x < i) | (x >= k), x, np.where(x < i + head, j + (x - i), x - head))
    return np.take_along_axis(cur, src, 1)


def sort_rows(slot: int, rows: np.ndarray, ns, max_ratio: float, max_depth: int, e0=None, norms=None,
              flags: int = 0) -> LadderResult:
    """sort_batch on one-line rows (count, 64) uint8 with pi the identity;
    e0 (odd cycles) and norms (3-norms) default to those of the rows."""
    count = len(ns)
    ns = np.asarray(ns, np.int32)
    rows = np.ascontiguousarray(rows, np.uint8)
    min_rate = float(np.float32(2.0) / np.float32(max_ratio))
    e0 = odd_cycle_counts(rows, ns) if e0 is None else np.asarray(e0, np.int64)
    norms = (ns - e0) // 2 if norms is None else np.asarray(norms, np.int64)
    # SortOrExtend.java:91-105 in float32: ub = floor(3-norm * (2 / minRate))
    ub = np.floor(norms.astype(np.float32) * (np.float32(2.0) / np.float32(min_rate)))
    rungs = np.stack([np.full(count, 4.0), np.minimum(ub, 8), np.minimum(ub, max_depth)], 1).astype(np.int32)
    nrungs = 1 + (ub > 4) + (ub > 8)
    ident = np.ascontiguousarray(np.broadcast_to(np.arange(64, dtype=np.uint8), (count, 64)))
    found = np.zeros(count, bool)
    tried = np.zeros(count, np.int32)
    moves: List[Optional[List[Move]]] = [None] * count
    launches = 0

    def convert(idx: np.ndarray, depth: np.ndarray, pos_all: np.ndarray, width_all: np.ndarray) -> None:
        """Positions -> symbols, replayed on pi (the identity) for all found
        configurations of each length at once (GPUSortingSearch.java:65-87)."""
        for d in np.unique(depth).tolist():
            sel = depth == d
            pos = pos_all[sel, :d].astype(np.int64)
            # only the first max n positions can move
            width = int(width_all[sel].max())
            cur = np.ascontiguousarray(np.broadcast_to(np.arange(width, dtype=np.uint8), (len(pos), width)))
            sym = np.empty(pos.shape, np.uint8)
            for t in range(d):
                sym[:, t] = np.take_along_axis(cur, pos[:, t], 1)
                cur = _block_swap_rows(cur, pos[:, t, 0], pos[:, t, 1], pos[:, t, 2])
            # one (a, b, c) tuple per move straight from the three symbol
            # columns, then each configuration's d consecutive moves
            flat = sym.reshape(-1, 3)
            tuples = list(zip(flat[:, 0].tolist(), flat[:, 1].tolist(), flat[:, 2].tolist()))
            for c, lo in zip(idx[sel].tolist(), range(0, len(tuples), d)):
                moves[c] = tuples[lo:lo + d]

    # a rung's sequences are converted on a host thread while the GPU runs the
    # next rung (the library call releases the interpreter lock); the thread
    # gets copies, since the next call reuses the result buffer
    with ThreadPoolExecutor(1) as converter:
        pending = []
        for rung in range(3):
            todo = np.flatnonzero(~found & (rung < nrungs))
            if todo.size == 0:
                break
            res = search_batch_rows(slot, ident[todo], rows[todo], ns[todo], e0[todo], min_rate, rungs[todo, rung],
                                    0, flags=flags, move_lists=False)
            launches += 1
            tried[todo] = rungs[todo, rung]
            hit = np.flatnonzero(res.found)
            found[todo[hit]] = True
            pending.append(converter.submit(convert, todo[hit], np.asarray(res.depth)[hit],
                                            np.asarray(res.moves_array)[hit], ns[todo[hit]]))
        for job in pending:
            job.result()
    return LadderResult(found, tried, moves, launches)


def write_cases(path: str, configs: Sequence[Cycles], result: LadderResult) -> int:
    """Writes the configurations sort_batch found sequences for as a case file
    (``spi->[(a b c), ...]``, the format of the shipped cases-*.txt that
    AbstractSbtAlgorithm.loadSortings reads); returns the number of lines."""
    from .case_files import from_results, write_case_file
    lines = from_results(configs, result.moves)
    write_case_file(path, lines)
    return len(lines)
