"""GPU parity tests of the search path, through the C ABI.

Deterministic modes (one wave per configuration) must match the CPU oracle
bit-for-bit: verdict, witness positions and every counter.  The parallel mode
(work sharing between waves) must match verdicts, and every witness must
replay to an accepted state.
"""
import threading
import time

import numpy as np
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import workloads as W
from helpers import MIN_RATE_15, MIN_RATE_1375, assert_witness, case_inputs, load_cases

pytestmark = pytest.mark.gpu

ORACLE_COUNTERS = ("expansions", "generated", "admitted", "dominated", "improved", "inserted", "roots_live")


def oracle_all(cases, min_rate, max_depth, full=False):
    res, _ = oracle.search_many([c.pi for c in cases], [c.spi for c in cases], [c.n for c in cases],
                                [c.even_cycles for c in cases], [min_rate] * len(cases),
                                [max_depth] * len(cases), [int(full)] * len(cases), True, 0, 8)
    return res


def gpu_batch(slot, cases, min_rate, max_depth, flags, full=False):
    return T.search_batch(slot, [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases],
                          min_rate, max_depth, int(full), flags=flags)


def assert_batch_parity(cases, got, want, counters=True):
    for c, case in enumerate(cases):
        w = want[c]
        assert bool(got.found[c]) == w.found, f"case {c}: verdict"
        if counters:
            assert int(got.depth[c]) == w.depth, f"case {c}: depth"
            assert [tuple(m) for m in got.moves[c]] == [tuple(m) for m in w.moves], f"case {c}: witness"
            for name in ORACLE_COUNTERS:
                assert int(got.counters[name][c]) == w.counters[name], f"case {c}: counter {name}"
            assert int(got.counters["uninserted"][c]) == 0


def test_eval_children_matches_recount(gpu_lib):
    """O(1) parity delta on the GPU == O(n) recount, every triple."""
    rng = np.random.default_rng(7)
    for it in range(120):
        n = int(rng.integers(3, 64))
        pi = [int(x) for x in rng.permutation(n)]
        if it % 3 == 0:   # sparse cycle structure: mostly fixed points
            cyc = []
            sym = [int(x) for x in rng.permutation(n)]
            p = 0
            while p < n // 2:
                L = int(rng.integers(1, 8))
                cyc.append(sym[p:p + L])
                p += L
            spi = W.one_line(cyc, n)
        else:
            spi = [int(x) for x in rng.permutation(n)]
        want = oracle.children_even(pi, spi)
        got = np.zeros(len(want), np.int32)
        import ctypes
        T._lib.check(gpu_lib.tdp_eval_children(T._lib.u8(pi, 64), T._lib.u8(spi, 64), n,
                                               got.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        assert np.array_equal(got, want), f"n={n} it={it}"


def _plus2_stage(pi, spi, e, goal, admit, node):
    """The cycle-local enumeration's result for one node from the oracle's
    recount: the largest +2 triple index if +2 children are goals, else the
    admitted +2 children in descending triple order as stage entries."""
    n = len(pi)
    odd = oracle.children_even(pi, spi)
    triples = [(i, j, k) for i in range(n) for j in range(i + 1, n) for k in range(j + 1, n)]
    plus2 = [t for t in range(len(triples)) if odd[t] == e + 2]
    if goal and plus2:
        return max(plus2), []
    if not admit:
        return -1, []
    return -1, [i | (j << 6) | (k << 12) | ((e + 2) << 18) | (node << 24)
                for (i, j, k) in (triples[t] for t in sorted(plus2, reverse=True))]


def test_paired_enumeration_matches_sequential_and_recount(gpu_lib):
    """Two batch nodes enumerated in one pass (n <= 32, +2-only) give the
    staged children and goal of one pass per node, and both equal the
    oracle's recount."""
    import ctypes
    rng = np.random.default_rng(23)
    rec = 4 + 512
    paired = 0
    for it in range(150):
        n = int(rng.integers(3, 33))
        nodes = []
        for _ in range(2):
            sym = [int(x) for x in rng.permutation(n)]
            cyc, p = [], 0
            even = it % 10 == 9
            while p < n:
                lengths = [x for x in ([1, 3, 5, 7] if not even else [1, 2, 3, 4, 5]) if x <= n - p]
                L = int(rng.choice(lengths))
                cyc.append(sym[p:p + L])
                p += L
            spi = W.one_line(cyc, n)
            pi = [int(x) for x in rng.permutation(n)]
            nodes.append((pi, spi, W.odd_cycle_count(spi)))
        (pa, sa, ea), (pb, sb, eb) = nodes
        ga, aa, gb, ab = [4 * int(x) for x in rng.integers(0, 2, 4)]
        aa = 4 if ga == 0 else aa
        ab = 4 if gb == 0 else ab
        out = np.zeros(2 * rec, np.int32)
        T._lib.check(gpu_lib.tdp_eval_pair(T._lib.u8(pa, 64), T._lib.u8(sa, 64), T._lib.u8(pb, 64),
                                           T._lib.u8(sb, 64), n, ea, eb, ga, aa, gb, ab,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        pr, sq = out[:rec], out[rec:]
        if pr[0] != 1:
            continue   # the pair declined (an even cycle, or too many candidates): one node at a time
        paired += 1
        assert sq[0] == 1, f"it={it}"
        assert list(pr[1:3]) == list(sq[1:3]), f"it={it} n={n}"
        if pr[1] < 0:   # no goal: the same children staged in the same order
            assert pr[3] == sq[3] and list(pr[4:4 + pr[3]]) == list(sq[4:4 + sq[3]]), f"it={it} n={n}"
        # the oracle: node a (batch index 1) first, then b (index 0)
        ta, sta = _plus2_stage(pa, sa, ea, ga, aa, 1)
        if ta >= 0:
            want = (1, ta, [])
        else:
            tb, stb = _plus2_stage(pb, sb, eb, gb, ab, 0)
            want = (0, tb, []) if tb >= 0 else (-1, -1, sta + stb)
        got_stage = [] if want[0] >= 0 else [int(x) for x in pr[4:4 + pr[3]]]
        assert (int(pr[1]), int(pr[2]), got_stage) == want, f"it={it} n={n}"
    assert paired >= 90


def test_single_wave_parity_random_n12(slots):
    """One wave, one configuration at a time: the reference's Q=1 order."""
    wl = W.workload_random_n12()
    cases = wl.cases[:48]
    want = oracle_all(cases, wl.min_rate, wl.max_depth)
    got = gpu_batch(slots[0], cases, wl.min_rate, wl.max_depth, T.FLAG_SINGLE_WAVE)
    assert_batch_parity(cases, got, want)


def test_deterministic_batch_parity_random_n12(slots):
    """512 x n=12, D=6 in one launch, one wave per configuration."""
    wl = W.workload_random_n12()
    want = oracle_all(wl.cases, wl.min_rate, wl.max_depth)
    got = gpu_batch(slots[0], wl.cases, wl.min_rate, wl.max_depth, T.FLAG_DETERMINISTIC)
    assert_batch_parity(wl.cases, got, want)


def test_deterministic_batch_parity_four_3cycles(slots):
    """n=24, D=9, both the oriented (as worded) and unoriented variants."""
    for oriented in (True, False):
        wl = W.workload_four_3cycles_n24(oriented)
        cases = wl.cases[:96]
        want = oracle_all(cases, wl.min_rate, wl.max_depth)
        got = gpu_batch(slots[1], cases, wl.min_rate, wl.max_depth, T.FLAG_DETERMINISTIC)
        assert_batch_parity(cases, got, want)


def test_parallel_batch_verdicts_random_n12(slots):
    wl = W.workload_random_n12()
    want = oracle_all(wl.cases, wl.min_rate, wl.max_depth)
    got = gpu_batch(slots[0], wl.cases, wl.min_rate, wl.max_depth, 0)
    assert_batch_parity(wl.cases, got, want, counters=False)
    for c, case in enumerate(wl.cases):
        if got.found[c]:
            assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])


def test_parallel_single_searches_match_oracle(slots):
    """The reference's one-search-per-call API with work sharing across waves."""
    wl = W.workload_random_n12()
    gs = T.GPUSortingSearch.__new__(T.GPUSortingSearch)
    for case in wl.cases[:24]:
        want = oracle.search(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth)
        res = T.search_for_sorting_seq(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, False,
                                       slots[0])
        assert bool(res.positions) == want.found
        if res.positions:
            assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, res.positions)
    del gs


def test_case_files_found(slots):
    """Every configuration of the reference's case files is FOUND at D = witness length."""
    data = load_cases()
    comb4 = [e for e in data["comb"] if len(e["witness"]) == 4]
    comb_deep = [e for e in data["comb"] if len(e["witness"]) > 4][::12]
    groups = (
        (data["3_2"], MIN_RATE_15, (T.FLAG_DETERMINISTIC, 0)),
        (data["oriented_7cycle"], MIN_RATE_1375, (T.FLAG_DETERMINISTIC, 0)),
        (comb4, MIN_RATE_1375, (T.FLAG_DETERMINISTIC, 0)),
        # (8,12) and (11,16) entries: the reference's depth-first order needs up
        # to millions of expansions on one wave; the work-sharing mode is used
        (comb_deep, MIN_RATE_1375, (0,)),
    )
    for entries, rate, modes in groups:
        pis, spis, e0s, ds = [], [], [], []
        for ent in entries:
            pi, spi = case_inputs(ent)
            pis.append(pi)
            spis.append(spi)
            e0s.append(W.odd_cycle_count(spi))
            ds.append(len(ent["witness"]))
        for flags in modes:
            got = T.search_batch(slots[1], pis, spis, e0s, rate, ds, 0, flags=flags)
            assert got.found.all(), (flags, np.nonzero(~got.found))
            for c in range(len(pis)):
                assert_witness(pis[c], spis[c], e0s[c], rate, ds[c], got.moves[c])


def test_case_files_deterministic_parity_small(slots):
    """Counters and witnesses on the small case-file configurations (ratio 1.5)."""
    data = load_cases()
    cases = [W.Case(*case_inputs(e)) for e in data["3_2"]]
    want = oracle_all(cases, MIN_RATE_15, 4)
    got = gpu_batch(slots[0], cases, MIN_RATE_15, 4, T.FLAG_DETERMINISTIC)
    assert_batch_parity(cases, got, want)


def test_bad_small_components_not_found(slots):
    """The six bad small components admit no (11/8)-sequence (Combinations.java:31-46)."""
    cases = W.bad_small_components()
    want = oracle_all(cases, MIN_RATE_1375, 11)
    assert not any(w.found for w in want)
    for flags in (T.FLAG_DETERMINISTIC, 0):
        got = gpu_batch(slots[0], cases, MIN_RATE_1375, 11, flags)
        assert not got.found.any()
    got = gpu_batch(slots[0], cases, MIN_RATE_1375, 11, T.FLAG_DETERMINISTIC)
    assert_batch_parity(cases, got, want)


def test_full_sorting_verdicts(slots):
    """fullSorting mode: GPU == oracle (same semantics) == brute force == SortingSearch.java restatement."""
    rng = W.SplitMix64(11)
    cases = []
    for _ in range(40):
        n = 5 + rng.below(4)
        pi = [0] + rng.shuffle(list(range(1, n)))
        cases.append(W.Case(pi, W.spi_from_pi(pi)))
    rate = MIN_RATE_15
    for D in (3, 5, 7):
        want = oracle_all(cases, rate, D, full=True)
        got = gpu_batch(slots[0], cases, rate, D, T.FLAG_DETERMINISTIC, full=True)
        assert_batch_parity(cases, got, want)
        for c, case in enumerate(cases):
            bfs = oracle.bfs_verdict(case.pi, case.spi, case.even_cycles, rate, D, full=True)
            java_f32 = oracle.java_full_sort_search(case.pi, case.spi, case.even_cycles, rate, D,
                                                    float_prune=True)
            java = oracle.java_full_sort_search(case.pi, case.spi, case.even_cycles, rate, D)
            assert bool(got.found[c]) == bfs == bool(java_f32)
            assert not java or got.found[c]


def test_brute_force_verdicts_small(slots):
    """Partial mode verdicts == exhaustive BFS over spi states (n <= 9)."""
    rng = W.SplitMix64(5)
    cases = []
    for _ in range(64):
        n = 6 + rng.below(4)
        pi = [0] + rng.shuffle(list(range(1, n)))
        cases.append(W.Case(pi, W.spi_from_pi(pi)))
    for rate, D in ((MIN_RATE_1375, 4), (MIN_RATE_1375, 6), (MIN_RATE_15, 3)):
        got = gpu_batch(slots[1], cases, rate, D, 0)
        for c, case in enumerate(cases):
            assert bool(got.found[c]) == oracle.bfs_verdict(case.pi, case.spi, case.even_cycles, rate, D)


def test_invalid_arguments(slots):
    pi = list(range(12))
    spi = W.spi_from_pi(pi)
    with pytest.raises(T.TdpError, match="n exceeds MAX_N"):
        T.search_for_sorting_seq(list(range(64)), list(range(64)), 64, 1.0, 4, False, slots[0])
    for D in (0, 16):
        with pytest.raises(T.TdpError, match="invalid maximumDepth"):
            T.search_for_sorting_seq(pi, spi, 12, 1.0, D, False, slots[0])
    with pytest.raises(T.TdpError, match="invalid slot"):
        T.search_for_sorting_seq(pi, spi, 12, 1.0, 4, False, 7)
    # n < 3: no triples, empty result (sorting_search.cu:718-719)
    assert T.search_for_sorting_seq([0, 1], [1, 0], 0, 1.0, 4, False, slots[0]).positions == []


def test_largest_size_and_empty_batch(slots):
    """n = 63 (MAX_N): 21 oriented 3-cycles give a 2-move at depth 1; random
    n = 63 configurations at depth 2 (39,711 children per node) return
    witnesses that replay; an empty batch returns empty results."""
    empty = T.search_batch(slots[0], [], [], [], MIN_RATE_1375, 4, 0)
    assert len(empty.found) == 0 and empty.moves == []
    n = 63
    pi = list(range(n))
    spi = [3 * (x // 3) + (x % 3 + 1) % 3 for x in range(n)]          # (0 1 2)(3 4 5)...
    e0 = W.odd_cycle_count(spi)
    r = T.search_for_sorting_seq(pi, spi, e0, MIN_RATE_1375, 4, False, slots[0])
    assert len(r.positions) == 1
    assert_witness(pi, spi, e0, MIN_RATE_1375, 4, r.positions)
    rng = W.SplitMix64(63)
    cases = []
    for _ in range(8):
        p = [0] + rng.shuffle(list(range(1, n)))
        cases.append(W.Case(p, W.spi_from_pi(p)))
    got = gpu_batch(slots[0], cases, MIN_RATE_1375, 2, 0)
    for c, case in enumerate(cases):
        if got.found[c]:
            assert_witness(case.pi, case.spi, case.even_cycles, MIN_RATE_1375, 2, got.moves[c])


def test_facade_returns_symbol_triples(slots):
    """GPUSortingSearch.search returns symbols replayed from positions (GPUSortingSearch.java:65-75)."""
    g = T.GPUSortingSearch.__new__(T.GPUSortingSearch)
    import queue
    g._slot_pool = queue.Queue()
    g._slot_pool.put(0)
    g.last_counters = None
    data = load_cases()
    for ent in data["3_2"][:4]:
        pi, spi = case_inputs(ent)
        e0 = W.odd_cycle_count(spi)
        moves = g.search(bytes(pi), spi, e0, MIN_RATE_15, 3, False)
        assert moves
        e, _, _ = oracle.replay_symbols(pi, spi, moves)
        assert oracle.satisfies_rate(e, len(moves), len(pi), e0, MIN_RATE_15, False)


def test_graph_rounds_match_eager(slots):
    """TDP_FLAG_GRAPH runs each launch round (reset kernel, search kernel,
    control-line copy) as the slot's HIP graph, captured and applied to the
    executable graph with hipGraphExecUpdate.  Deterministic mode: three graph
    calls, each followed by an eager call, all equal to the oracle bit for bit
    (verdicts, witnesses, counters).  Work-sharing mode over 4 rounds of 128
    configurations (max_active 128): graph and eager verdicts equal, every
    witness replays, on two graph calls with an eager call between."""
    wl = W.workload_random_n12()
    cases = wl.cases[:96]
    want = oracle_all(cases, wl.min_rate, wl.max_depth)
    det = T.FLAG_DETERMINISTIC
    for rep in range(3):
        got = gpu_batch(slots[0], cases, wl.min_rate, wl.max_depth, det | T._lib.FLAG_GRAPH)
        assert_batch_parity(cases, got, want)
        eager = gpu_batch(slots[0], cases, wl.min_rate, wl.max_depth, det)
        assert_batch_parity(cases, eager, want)

    batch = wl.cases
    T.set_tuning(max_active=128)
    try:
        eager = gpu_batch(slots[1], batch, wl.min_rate, wl.max_depth, 0)
        for rep in range(2):
            got = gpu_batch(slots[1], batch, wl.min_rate, wl.max_depth, T._lib.FLAG_GRAPH)
            assert np.array_equal(got.found, eager.found)
            for c, case in enumerate(batch):
                if got.found[c]:
                    assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])
            again = gpu_batch(slots[1], batch, wl.min_rate, wl.max_depth, 0)
            assert np.array_equal(again.found, eager.found)
    finally:
        T.set_tuning()
