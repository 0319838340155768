"""GPU results over whole workloads against the oracle's committed results
(tests/golden/make_search_verdicts.py writes them on the CPU):

* the 2,048 n = 24 configurations of both four-3-cycle variants (max_ratio
  1.375, D = 9, dedup 2^24, 3 GiB of queue memory, 2 slots): the
  deterministic mode reproduces the oracle's verdict, sequence and every
  counter of every configuration; the work-sharing mode its verdicts;
* the 4,096-configuration depth-11 batch exactly as bench.py runs it (each
  ordering with its own cached tuning, work-sharing, dedup 2^27, 32 GiB of
  queue memory), with the plain and with the learned ordering: every verdict
  the oracle decided within its expansion budget is equal.

Every returned sequence is replayed to an accepted state.
"""
import json
import os

import numpy as np
import pytest

import tdp1375_amd as T
from tdp1375_amd import autotune
from tdp1375_amd import workloads as W
from helpers import GOLDEN, assert_witness

pytestmark = pytest.mark.gpu


def _fixture(name):
    return np.load(os.path.join(GOLDEN, f"search_{name}.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def n24_slots(gpu_lib):
    T.init_slots(2, 1, 3072 << 20, 1 << 24)
    yield [0, 1]
    T.destroy_slots()


@pytest.mark.parametrize("oriented", [True, False])
def test_four_3cycles_deterministic_all_2048(n24_slots, oriented):
    wl = W.workload_four_3cycles_n24(oriented)
    fx = _fixture(wl.name)
    names = [str(x) for x in fx["counter_names"]]
    got = T.search_batch(n24_slots[1], [c.pi for c in wl.cases], [c.spi for c in wl.cases],
                         [c.even_cycles for c in wl.cases], wl.min_rate, wl.max_depth, 0,
                         flags=T.FLAG_DETERMINISTIC)
    assert T.last_launch_stats(n24_slots[1])["evictions"] == 0
    assert np.array_equal(got.found.astype(np.uint8), fx["found"])
    for c in range(len(wl.cases)):
        d = int(fx["depth"][c])
        assert int(got.depth[c]) == d, f"case {c}: depth"
        assert [tuple(m) for m in got.moves[c]] == [tuple(int(x) for x in fx["moves"][c, m]) for m in range(d)], \
            f"case {c}: sequence"
    for k, name in enumerate(names):
        assert np.array_equal(got.counters[name].astype(np.uint64), fx["counters"][:, k]), name
    assert int(got.counters["uninserted"].sum()) == 0


@pytest.mark.parametrize("oriented", [True, False])
def test_four_3cycles_parallel_verdicts_all_2048(n24_slots, oriented):
    wl = W.workload_four_3cycles_n24(oriented)
    fx = _fixture(wl.name)
    got = T.search_batch(n24_slots[0], [c.pi for c in wl.cases], [c.spi for c in wl.cases],
                         [c.even_cycles for c in wl.cases], wl.min_rate, wl.max_depth, 0, flags=0)
    assert np.array_equal(got.found.astype(np.uint8), fx["found"])
    for c, case in enumerate(wl.cases):
        if got.found[c]:
            assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])


def _depth11_verdicts():
    path = os.path.join(GOLDEN, "verdicts_odd-cycle-mix-n9-36-d11.json")
    with open(path) as fh:
        return json.load(fh)


def run_depth11_bench_path(flags=0, cases=4096):
    """The bench's step: the whole batch (or its first `cases`
    configurations), tuned for its ordering, work-sharing mode."""
    from tdp1375_amd import _lib
    wl = W.workload_odd_cycle_mix_d11(cases)
    T.init_slots(1, 1, 32768 << 20, 1 << 27)
    try:
        autotune.apply(autotune.DEFAULT_CACHE,
                       autotune.cache_key(wl.name, "learned" if flags & _lib.FLAG_LEARNED else "plain"))
        got = T.search_batch(0, [c.pi for c in wl.cases], [c.spi for c in wl.cases],
                             [c.even_cycles for c in wl.cases], wl.min_rate, wl.max_depth, 0, flags=flags)
    finally:
        autotune.reset()
        T.destroy_slots()
    return wl, got


def check_depth11(wl, got):
    fx = _depth11_verdicts()
    verdicts = fx["verdicts"]
    assert len(verdicts) == len(wl.cases)
    decided = [c for c, v in enumerate(verdicts) if v != "?"]
    want = np.array([verdicts[c] == "1" for c in decided])
    assert np.array_equal(got.found[decided], want), \
        [c for c, w in zip(decided, want) if bool(got.found[c]) != w][:20]
    for c, case in enumerate(wl.cases):
        if got.found[c]:
            assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])
    return len(decided)


def test_depth11_batch_bench_path_verdicts(gpu_lib):
    wl, got = run_depth11_bench_path()
    decided = check_depth11(wl, got)
    assert decided >= 3900


def test_depth11_batch_learned_ordering_verdicts(gpu_lib):
    """The learned ordering (committed model) over the same batch and path:
    every decided verdict equals the oracle's and every sequence replays."""
    from tdp1375_amd import _lib, train_ordering
    train_ordering.load(os.path.join(os.path.dirname(T.__file__), "data", "ordering_mlp_d11.npz"))
    wl, got = run_depth11_bench_path(flags=_lib.FLAG_LEARNED)
    decided = check_depth11(wl, got)
    assert decided >= 3900


def test_trained_network_gains_over_null_network_and_plain(gpu_lib):
    """What the trained network gains, on the first 1,024 configurations of
    the depth-11 batch (bench path): the learned kernel with the committed
    model against the same kernel with an all-zero network (the same
    hold-back and push machinery, every child scored alike) and against the
    plain ordering.  Verdicts are equal in all three; expansions to verdict
    over the satisfiable configurations must be at least 3x fewer with the
    trained network than with either (measured on the whole batch: 3.2e7
    trained, 9.3e8 plain, 2.7e9 zero network)."""
    from tdp1375_amd import _lib, learned as LN, train_ordering
    model = os.path.join(os.path.dirname(T.__file__), "data", "ordering_mlp_d11.npz")
    _, plain = run_depth11_bench_path(cases=1024)
    with np.load(model, allow_pickle=False) as z:
        w1, b1, w2 = z["w1"], z["b1"], z["w2"]
    LN.set_weights(np.zeros_like(w1), np.zeros_like(b1), np.zeros_like(w2), 0.0)
    try:
        _, null = run_depth11_bench_path(flags=_lib.FLAG_LEARNED, cases=1024)
    finally:
        train_ordering.load(model)
    wl, trained = run_depth11_bench_path(flags=_lib.FLAG_LEARNED, cases=1024)
    assert np.array_equal(plain.found, null.found) and np.array_equal(plain.found, trained.found)
    verdicts = _depth11_verdicts()["verdicts"][:1024]
    decided = [c for c, v in enumerate(verdicts) if v != "?"]
    assert np.array_equal(trained.found[decided], np.array([verdicts[c] == "1" for c in decided]))
    sat = plain.found
    exp = {name: float(r.counters["expansions"][sat].sum()) for name, r in
           (("plain", plain), ("zero network", null), ("trained", trained))}
    print(f"satisfiable {int(sat.sum())} of {len(sat)}: expansions to verdict {exp}", flush=True)
    assert exp["plain"] >= 3.0 * exp["trained"]
    assert exp["zero network"] >= 3.0 * exp["trained"]
    for c, case in enumerate(wl.cases):
        if trained.found[c]:
            assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, trained.moves[c])
