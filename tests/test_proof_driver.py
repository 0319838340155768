"""The proof-generation driver: iterative deepening (SortOrExtend.java:89-111),
single flight (SortingCoordinator.java:72-98), resume files
(SortOrExtend.java:53-87), and the batched ladder on the GPU."""
import os
import threading

import numpy as np
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import proof_driver as PD
from tdp1375_amd import workloads as W
from helpers import MIN_RATE_15, MIN_RATE_1375, load_cases


def test_permutation_helpers_known_answers():
    cycles = [(3, 4, 0), (1, 5, 2), (6,)]
    assert PD.spi_string(cycles) == "(0 3 4)(1 5 2)(6)"
    assert PD.one_line(cycles) == [3, 5, 1, 4, 0, 2, 6]
    assert PD.odd_cycles(cycles) == 3
    assert PD.three_norm(cycles) == 2          # (7 - 3) / 2
    assert PD.spi_string([]) == "()"


def test_depth_ladder_matches_the_reference_rungs():
    # SortOrExtend.java:95-105: 4, then min(ub, 8) if ub > 4, then min(ub, D) if ub > 8
    assert PD.depth_ladder(2, 11) == [4]
    assert PD.depth_ladder(4, 11) == [4]
    assert PD.depth_ladder(6, 11) == [4, 6]
    assert PD.depth_ladder(8, 11) == [4, 8]
    assert PD.depth_ladder(9, 11) == [4, 8, 9]
    assert PD.depth_ladder(20, 11) == [4, 8, 11]
    # ub = floor(3-norm * 1.375) with the float32 quotient 2 / minRate
    eight_3cycles = [(3 * i, 3 * i + 2, 3 * i + 1) for i in range(8)]
    assert PD.three_norm(eight_3cycles) == 8
    assert PD.upper_bound(eight_3cycles, MIN_RATE_1375) == 11.0
    assert PD.upper_bound(eight_3cycles, MIN_RATE_15) == 12.0


def test_coordinator_single_flight():
    co = PD.SortingCoordinator()
    started, release = threading.Event(), threading.Event()
    out = {}

    def leader():
        def work():
            started.set()
            release.wait(5)
            return "sorting"
        out["leader"] = co.try_compute("k", work)

    th = threading.Thread(target=leader)
    th.start()
    assert started.wait(5)
    second = co.try_compute("k", lambda: "other")
    assert second.is_already_in_progress
    assert co.try_compute("other-key", lambda: None).is_not_found
    release.set()
    th.join()
    assert out["leader"].is_computed and out["leader"].value == "sorting"
    # the key is released: a later call computes again
    assert co.try_compute("k", lambda: "again").value == "again"
    with pytest.raises(ValueError):
        co.try_compute("k", lambda: (_ for _ in ()).throw(ValueError("boom")))
    assert co.try_compute("k", lambda: None).is_not_found


class OracleSearch:
    """Stands in for GPUSortingSearch on the CPU (same search() contract)."""

    def __init__(self):
        self.calls = []

    def search(self, pi, spi, e0, min_rate, max_moves, full_sorting, flags=0):
        self.calls.append(max_moves)
        r = oracle.search(pi, spi, e0, min_rate, max_moves, full_sorting)
        if not r.found:
            return []
        return T.positions_to_symbols(pi, r.moves)


def test_sort_or_extend_resume_and_extend(tmp_path):
    data = load_cases()
    good = [tuple(c) for c in data["3_2"][0]["spi"]]
    gpu = OracleSearch()
    extended = []
    drv = PD.SortOrExtend(gpu, 1.5, 11, str(tmp_path), extend=extended.append)
    r = drv.compute(good)
    assert r.is_computed and r.value is True and extended == []
    path = drv.sorting_path(PD.spi_string(good))
    moves = [tuple(int(x) for x in line.strip("()\n").split()) for line in open(path)]
    n = sum(len(c) for c in good)
    e, _, _ = oracle.replay_symbols(list(range(n)), PD.one_line(good), moves)
    assert oracle.satisfies_rate(e, len(moves), n, PD.odd_cycles(good), MIN_RATE_15, False)
    calls = len(gpu.calls)
    assert drv.compute(good).is_computed and len(gpu.calls) == calls      # resumed from the file

    # the interleaving pair admits no (11/8)-sequence (index.html:52-63):
    # searched once, marked as a bad case, extended; a second call only reads
    # the marker and extends again (SortOrExtend.java:62-85)
    # (its own output directory: at ratio 1.5 the same spi has a sequence)
    bad = list(W.BAD_SMALL_COMPONENTS["interleaving_pair"])
    out = str(tmp_path / "r1375")
    drv = PD.SortOrExtend(gpu, 1.375, 11, out, extend=extended.append)
    calls = len(gpu.calls)
    assert drv.compute(bad).is_not_found and extended == [bad]
    assert len(gpu.calls) > calls
    assert os.path.exists(os.path.join(out, "bad-cases", PD.spi_string(bad)))
    calls = len(gpu.calls)
    assert drv.compute(bad).is_not_found and len(gpu.calls) == calls
    assert extended == [bad, bad]


def test_search_for_sorting_walks_the_ladder():
    """A comb (8,12) entry has no 4-move sequence and one at depth 8: the
    driver tries D = 4, then D = 8 (ub = 11 > 8 stops there once found)."""
    data = load_cases()
    entry = next(e for e in data["comb"] if len(e["witness"]) == 8 and sum(len(c) for c in e["spi"]) <= 24)
    cycles = [tuple(c) for c in entry["spi"]]
    gpu = OracleSearch()
    drv = PD.SortOrExtend(gpu, 1.375, 11, "/tmp/tdp_driver_unused")
    moves = drv.search_for_sorting(cycles)
    assert moves is not None and len(moves) <= 8
    assert gpu.calls[0] == 4 and gpu.calls[-1] == 8


def test_odd_cycle_counts_by_walking_every_cycle():
    """odd_cycle_counts (pointer doubling over the widest row's columns) on
    random one-line rows of n = 1..64, mixed and equal widths, against
    walking each cycle; fixed points are odd cycles."""
    rng = np.random.default_rng(7)
    for trial in range(60):
        K = int(rng.integers(1, 40))
        ns = rng.integers(1, 65, K) if trial % 2 else np.full(K, int(rng.integers(1, 12)))
        rows = np.tile(np.arange(64, dtype=np.uint8), (K, 1))
        want = []
        for c in range(K):
            rows[c, :ns[c]] = rng.permutation(int(ns[c]))
            seen, odd = set(), 0
            for s0 in range(int(ns[c])):
                length, x = 0, s0
                while x not in seen:
                    seen.add(x)
                    x = int(rows[c, x])
                    length += 1
                odd += length % 2
            want.append(odd)
        assert PD.odd_cycle_counts(rows, ns).tolist() == want, trial
    assert PD.odd_cycle_counts(np.zeros((0, 64), np.uint8), np.zeros(0, np.int32)).size == 0


def test_batched_ladder_on_the_oracle_equals_the_sequential_ladder(monkeypatch):
    """sort_batch (rows, one launch per rung, vectorised odd-cycle counts,
    bounds and positions -> symbols) with the oracle standing in for the
    batched search: the same verdicts, rungs and symbol moves as
    SortOrExtend.searchForSorting run one configuration at a time."""
    from tdp1375_amd.gpu_sorting_search import BatchResult
    calls = []

    def fake_rows(slot, P, S, ns, e0, min_rate, depths, full, flags=0, move_lists=True):
        calls.append(len(ns))
        res = [oracle.search(list(P[c, :ns[c]]), list(S[c, :ns[c]]), int(e0[c]), min_rate, int(depths[c]))
               for c in range(len(ns))]
        arr = np.zeros((len(ns), 15, 3), np.int64)
        for c, r in enumerate(res):
            if r.moves:
                arr[c, :len(r.moves)] = np.asarray(r.moves)
        return BatchResult(np.array([r.found for r in res]), np.array([r.depth for r in res], np.int32),
                           [r.moves for r in res] if move_lists else None, {}, arr)

    monkeypatch.setattr(PD, "search_batch_rows", fake_rows)
    data = load_cases()
    small = sorted(data["comb"], key=lambda e: sum(len(c) for c in e["spi"]))[:24]
    # configurations without an 11/8-sequence take every rung of their ladder
    # (index.html:44-59): depths 4 and 5 for n = 12, 4 and 6 for n = 15
    bad = ["(0 9 2)(1 11 4)(3 8 6)(5 10 7)", "(0 10 2)(1 5 3)(4 8 6)(7 11 9)", "(0 7 5)(1 11 9)(2 6 4)(3 10 8)",
           "(0 4 2)(1 14 12)(3 7 5)(6 10 8)(9 13 11)"]
    small = [{"spi": [[int(x) for x in c.split()] for c in b.strip("()").split(")(")]} for b in bad] + small
    for entries, ratio in ((data["3_2"], 1.5), (small, 1.375)):
        configs = [[tuple(c) for c in e["spi"]] for e in entries]
        res = PD.sort_batch(0, configs, ratio, 11)
        for c, cfg in enumerate(configs):
            gpu = OracleSearch()
            moves = PD.SortOrExtend(gpu, ratio, 11, "/tmp/tdp_driver_unused").search_for_sorting(cfg)
            assert bool(res.found[c]) == (moves is not None), c
            assert res.depth_tried[c] == gpu.calls[-1], c
            assert res.moves[c] == (moves if moves is not None else None), c
    assert calls and max(calls) <= 28


@pytest.mark.gpu
def test_batched_ladder_finds_every_case_file_configuration(gpu_lib):
    """Every configuration of cases-comb.txt (ratio 1.375) and cases-3,2.txt
    (ratio 1.5) is found by the batched ladder, each sequence replays, and no
    configuration needs a rung deeper than its witness."""
    data = load_cases()
    T.init_slots(1, 1, 2 << 30, 1 << 22)
    try:
        for entries, ratio, rate in ((data["comb"], 1.375, MIN_RATE_1375), (data["3_2"], 1.5, MIN_RATE_15)):
            configs = [[tuple(c) for c in e["spi"]] for e in entries]
            res = PD.sort_batch(0, configs, ratio, 11)
            print(f"ratio {ratio}: {int(res.found.sum())}/{len(configs)} found in {res.launches} launches",
                  flush=True)
            assert res.found.all()
            for c, e in enumerate(entries):
                n = sum(len(x) for x in configs[c])
                moves = res.moves[c]
                assert 1 <= len(moves) <= res.depth_tried[c]
                assert res.depth_tried[c] <= max(4, len(e["witness"]))
                ev, _, _ = oracle.replay_symbols(list(range(n)), PD.one_line(configs[c]), moves)
                assert oracle.satisfies_rate(ev, len(moves), n, PD.odd_cycles(configs[c]), rate, False)
    finally:
        T.destroy_slots()
