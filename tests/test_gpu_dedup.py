"""Dedup table under pressure: full lines evict entries, and verdicts do not
depend on what the table holds.

A dominated child is pruned only when a shallower-or-equal copy of the same
state was admitted before, so losing entries costs repeated work but never
changes whether a sequence is found.  These tests run the same batches with a
table far too small for them (constant evictions) and without a table, and
compare verdicts with the CPU oracle's exact dedup.
"""
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import workloads as W
from helpers import assert_witness

pytestmark = pytest.mark.gpu


def _oracle_verdicts(cases, min_rate, max_depth):
    res, _ = oracle.search_many([c.pi for c in cases], [c.spi for c in cases], [c.n for c in cases],
                                [c.even_cycles for c in cases], [min_rate] * len(cases),
                                [max_depth] * len(cases), [0] * len(cases), True, 0, 8)
    return [r.found for r in res]


@pytest.fixture(scope="module")
def small_tables(gpu_lib):
    yield
    T.destroy_slots()


@pytest.mark.parametrize("dedup", [64, 5, 0])
def test_small_table_keeps_verdicts(small_tables, dedup):
    wl = W.workload_random_n12()
    cases = wl.cases[:256]
    want = _oracle_verdicts(cases, wl.min_rate, wl.max_depth)
    T.init_slots(1, 1, 1 << 30, dedup)
    for flags in (T.FLAG_DETERMINISTIC, 0):
        got = T.search_batch(0, [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases],
                             wl.min_rate, wl.max_depth, 0, flags=flags)
        st = T.last_launch_stats(0)
        print(f"dedup {dedup} flags {flags}: evictions {st['evictions']} "
              f"uninserted {int(got.counters['uninserted'].sum())}", flush=True)
        assert [bool(f) for f in got.found] == want, (dedup, flags)
        for c, case in enumerate(cases):
            if got.found[c]:
                assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])
        if dedup == 64:
            assert st["evictions"] > 0
        if dedup == 0:
            assert st["evictions"] == 0 and int(got.counters["inserted"].sum()) == 0


def test_large_table_has_no_evictions(small_tables):
    """With room to spare no line fills: the table is exact (deterministic
    counters equal the oracle's, which the parity tests check)."""
    wl = W.workload_random_n12()
    T.init_slots(1, 1, 1 << 30, 1 << 22)
    got = T.search_batch(0, [c.pi for c in wl.cases], [c.spi for c in wl.cases], [c.even_cycles for c in wl.cases],
                         wl.min_rate, wl.max_depth, 0, flags=T.FLAG_DETERMINISTIC)
    st = T.last_launch_stats(0)
    assert st["evictions"] == 0
    assert int(got.counters["uninserted"].sum()) == 0
