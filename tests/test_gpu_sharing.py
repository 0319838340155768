"""Work sharing between waves: inbox hand-offs and the spill pool.

Small stacks force full stacks to hand nodes to idle waves or spill them to
the pool; verdicts must still equal the CPU oracle's and every witness must
replay.  The launch must also end by itself (no watchdog) however the work
moved.
"""
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import workloads as W
from helpers import assert_witness

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tuned(gpu_lib):
    yield
    T.set_tuning()
    T.destroy_slots()


def _verdicts(cases, wl):
    res, _ = oracle.search_many([c.pi for c in cases], [c.spi for c in cases], [c.n for c in cases],
                                [c.even_cycles for c in cases], [wl.min_rate] * len(cases),
                                [wl.max_depth] * len(cases), [0] * len(cases), True, 0, 8)
    return [r.found for r in res]


@pytest.mark.parametrize("waves_per_cu,budget", [(1, 2 << 20), (2, 3 << 20), (0, 64 << 20)])
def test_small_stacks_spill_and_share(tuned, waves_per_cu, budget):
    wl = W.workload_random_n12()
    cases = wl.cases[:96]
    want = _verdicts(cases, wl)
    T.set_tuning(waves_per_cu=waves_per_cu, share_interval=4, share_max=8)
    T.init_slots(1, 1, budget, 1 << 20)
    got = T.search_batch(0, [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases],
                         wl.min_rate, wl.max_depth, 0, flags=0)
    st = T.last_launch_stats(0)
    print(f"waves {st['waves']} budget {budget}: donations {st['donations']} spills {st['spills']} "
          f"takes {st['spill_takes']}", flush=True)
    assert [bool(f) for f in got.found] == want
    for c, case in enumerate(cases):
        if got.found[c]:
            assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])
    assert st["steals"] <= st["donations"]
    assert st["spill_takes"] <= st["spills"]
    if waves_per_cu == 1:
        # 220 root children against stacks of 64 nodes
        assert st["spills"] + st["donations"] > 0


def test_single_configuration_spreads_over_waves(tuned):
    """One configuration alone on the GPU: idle waves receive its nodes.  The
    configuration (depth-11 batch, index 20, n=26) has a sequence: the CPU
    oracle finds one after 62,184 expansions in the reference's order."""
    T.set_tuning()
    T.init_slots(1, 1, 1 << 30, 1 << 22)
    wl = W.workload_odd_cycle_mix_d11(64)
    case = wl.cases[20]
    got = T.search_batch(0, [case.pi], [case.spi], [case.even_cycles], wl.min_rate, wl.max_depth, 0, flags=0)
    st = T.last_launch_stats(0)
    assert st["donations"] > 0 and st["steals"] <= st["donations"]
    assert got.found[0]
    assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[0])


def test_batch_runs_in_rounds_of_max_active(tuned):
    """A batch larger than max_active runs as consecutive launches (fresh dedup
    epoch each); verdicts and witnesses are those of the whole batch, also when
    the configurations are re-run from device memory."""
    wl = W.workload_random_n12()
    cases = wl.cases[:200]
    want = _verdicts(cases, wl)
    T.set_tuning(max_active=48)
    T.init_slots(1, 1, 1 << 30, 1 << 20)
    args = ([c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases], wl.min_rate,
            wl.max_depth, 0)
    for flags in (0, T._lib.FLAG_RESIDENT):
        got = T.search_batch(0, *args, flags=flags)
        assert [bool(f) for f in got.found] == want
        for c, case in enumerate(cases):
            if got.found[c]:
                assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])
    T.set_tuning()


@pytest.mark.parametrize("nodes", [1, 3, 8])
def test_expansion_batches_keep_verdicts(tuned, nodes):
    """Expansion batches of 1, 3 and 8 nodes (one dedup pass over the children
    of several expanded nodes): verdicts equal the oracle's on the n=12 batch,
    also with small stacks that hand nodes over and spill, and on 256
    configurations of the depth-11 generator they equal those of one node at
    a time; every witness replays."""
    T.set_tuning()
    T.set_expansion_batch(nodes)
    try:
        wl = W.workload_random_n12()
        cases = wl.cases[:160]
        want = _verdicts(cases, wl)
        for budget in (3 << 20, 1 << 30):
            T.init_slots(1, 1, budget, 1 << 20)
            got = T.search_batch(0, [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases],
                                 wl.min_rate, wl.max_depth, 0, flags=0)
            assert [bool(f) for f in got.found] == want, budget
            for c, case in enumerate(cases):
                if got.found[c]:
                    assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])
        d11 = W.workload_odd_cycle_mix_d11(256)
        T.init_slots(1, 1, 4 << 30, 1 << 24)
        args = ([c.pi for c in d11.cases], [c.spi for c in d11.cases], [c.even_cycles for c in d11.cases],
                d11.min_rate, d11.max_depth, 0)
        par = T.search_batch(0, *args, flags=0)
        T.set_expansion_batch(1)
        one = T.search_batch(0, *args, flags=0)
        assert list(par.found) == list(one.found)
        for c, case in enumerate(d11.cases):
            if par.found[c]:
                assert_witness(case.pi, case.spi, case.even_cycles, d11.min_rate, d11.max_depth, par.moves[c])
    finally:
        T.set_expansion_batch(0)
