"""Learned move ordering: features, the MLP's forward/backward kernels against
the CPU reference, training on the GPU, and the search with the learned
ordering (same verdicts as the plain search, fewer expansions to verdict)."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from oracle import mlp_ref
import tdp1375_amd as T
from tdp1375_amd import _lib
from tdp1375_amd import learned as LN
from tdp1375_amd import workloads as W
from helpers import MIN_RATE_15, MIN_RATE_1375, assert_witness, case_inputs, load_cases


def random_state(rng, n):
    pi = [int(x) for x in rng.permutation(n)]
    spi = [int(x) for x in rng.permutation(n)]
    return pi, spi


def random_rows(rng, N, G):
    rows = rng.integers(0, 2 ** 32, size=(N, 4), dtype=np.uint64).astype(np.uint32)
    cuts = np.sort(rng.choice(np.arange(1, N), size=G - 1, replace=False))
    starts = np.concatenate([[0], cuts, [N]]).astype(np.int32)
    labels = np.zeros(N, np.uint8)
    for g in range(G):
        labels[rng.integers(starts[g], starts[g + 1])] = 1
    return rows, labels, starts


def random_weights(rng):
    return (rng.normal(0, 0.2, (128, 256)).astype(np.float32), rng.normal(0, 0.1, 256).astype(np.float32),
            rng.normal(0, 0.2, 256).astype(np.float32), float(rng.normal(0, 0.1)))


# ---------------------------------------------------------------- CPU
def test_reference_gradients_match_finite_differences():
    """Pins the CPU reference: its analytic gradient (without bf16 rounding)
    equals central differences of its loss on every parameter block."""
    rng = np.random.default_rng(3)
    rows, labels, starts = random_rows(rng, 40, 7)
    w1, b1, w2, b2 = [np.asarray(x, np.float64) if not np.isscalar(x) else x for x in random_weights(rng)]
    _, g = mlp_ref.loss_and_gradients(rows, labels, starts, w1, b1, w2, b2, exact=True)
    eps = 1e-6

    def fd(block, idx):
        args = [w1.copy(), b1.copy(), w2.copy(), b2]
        if block == 3:
            args[3] = b2 + eps
            up = mlp_ref.loss(rows, labels, starts, *args, exact=True)
            args[3] = b2 - eps
            return (up - mlp_ref.loss(rows, labels, starts, *args, exact=True)) / (2 * eps)
        a = args[block]
        a.flat[idx] += eps
        up = mlp_ref.loss(rows, labels, starts, *args, exact=True)
        a.flat[idx] -= 2 * eps
        return (up - mlp_ref.loss(rows, labels, starts, *args, exact=True)) / (2 * eps)

    offsets = [0, 128 * 256, 128 * 256 + 256, 128 * 256 + 512]
    for block, count in ((0, 128 * 256), (1, 256), (2, 256), (3, 1)):
        for idx in rng.choice(count, size=min(count, 12), replace=False):
            assert abs(fd(block, idx) - g[offsets[block] + idx]) < 1e-6, (block, idx)


def test_numpy_deltas_equal_oracle_recount():
    rng = np.random.default_rng(9)
    for it in range(30):
        n = int(rng.integers(3, 30))
        pi, spi = random_state(rng, n)
        delta, feats, tr = LN.deltas_and_features(pi, spi, 0, 0, 11)
        want = np.asarray(oracle.children_even(pi, spi)) - LN.odd_cycles(spi)
        assert np.array_equal(delta, want), (it, n)
        assert feats.shape == (len(tr), 4)


def test_feature_groups_are_one_hot():
    rng = np.random.default_rng(4)
    pi, spi = random_state(rng, 20)
    _, feats, _ = LN.deltas_and_features(pi, spi, 3, 2, 11)
    bits = mlp_ref.unpack(feats).astype(bool)
    L = LN.FEATURE_LAYOUT
    for name, width in (("delta", 3), ("child_depth", 16), ("gain", 16), ("moves_left", 16), ("len_a", 16),
                        ("len_b", 16), ("len_c", 16), ("sharing", 5), ("n_quarter", 16)):
        assert (bits[:, L[name]:L[name] + width].sum(axis=1) == 1).all(), name
    # known answer: pi = identity, spi = (0 1 2)(3 4 5), move (0, 1, 2) splits
    # the 3-cycle in cycle order into three fixed points: delta +2
    delta, feats, tr = LN.deltas_and_features(list(range(6)), [1, 2, 0, 4, 5, 3], 2, 0, 4)
    t = [tuple(x) for x in tr.tolist()].index((0, 1, 2))
    b = mlp_ref.unpack(feats[t:t + 1])[0].astype(bool)
    assert delta[t] == 2 and b[L["delta"] + 2] and b[L["sharing"] + 4] and b[L["cycle_order"]]
    assert b[L["len_a"] + 2] and b[L["child_depth"] + 1] and b[L["gain"] + 2] and b[L["moves_left"] + 3]


def test_dataset_groups_hold_one_taken_move():
    data = load_cases()
    cases, witnesses = [], []
    for e in data["3_2"]:
        pi, spi = case_inputs(e)
        r = oracle.search(pi, spi, W.odd_cycle_count(spi), MIN_RATE_15, 4)
        assert r.found
        cases.append(W.Case(pi, spi))
        witnesses.append(r.moves)
    ds = LN.build_dataset(cases, MIN_RATE_15, 4, witnesses)
    assert ds.groups > 0
    for g in range(ds.groups):
        s, e = ds.group_start[g], ds.group_start[g + 1]
        assert e - s > 1 and ds.labels[s:e].sum() == 1


# ---------------------------------------------------------------- GPU
@pytest.mark.parametrize("argv, message", [
    (["--seeds", "3,11"], "seed 3"),
    (["--validation-seed", "3"], "seed 3"),
    (["--validation-seed", "12"], "validation seed"),
    (["--candidates", "0"], "--candidates"),
])
def test_training_refuses_benchmark_and_overlapping_seeds(argv, message):
    # the benchmark batch (seed 3) never trains or picks a network, and the
    # validation batch is not a training batch; refused before any GPU call
    from tdp1375_amd import train_ordering
    with pytest.raises(SystemExit) as ex:
        train_ordering.main(["--out", "/nonexistent/model.npz"] + argv)
    assert message in str(ex.value)


@pytest.mark.gpu
def test_gpu_features_equal_numpy(gpu_lib):
    rng = np.random.default_rng(17)
    for it in range(40):
        n = int(rng.integers(3, 64))
        pi, spi = random_state(rng, n)
        e0, d, D = int(rng.integers(0, n)), int(rng.integers(0, 14)), int(rng.integers(1, 16))
        _, want, _ = LN.deltas_and_features(pi, spi, e0, d, D)
        got = np.zeros((len(want), 4), np.uint32)
        _lib.check(gpu_lib.tdp_eval_features(_lib.u8(pi, 64), _lib.u8(spi, 64), n, e0, d, D,
                                             got.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        assert np.array_equal(got, want), (it, n)


@pytest.mark.gpu
def test_forward_and_gradients_match_reference(gpu_lib):
    """MFMA forward and backward against the float64 reference with the same
    bf16 roundings; tolerances: logits 1e-4 relative to their scale, flat
    gradient 2e-3 relative in norm, b1/w2/b2 blocks 1e-3."""
    rng = np.random.default_rng(23)
    w1, b1, w2, b2 = random_weights(rng)
    LN.set_weights(w1, b1, w2, b2)
    for N, G in ((37, 5), (700, 40), (5000, 300)):
        rows, labels, starts = random_rows(rng, N, G)
        _, ref_logits = mlp_ref.forward(rows, w1, b1, w2, b2)
        got = LN.forward(rows)
        assert np.max(np.abs(got - ref_logits)) <= 1e-4 * max(1.0, np.max(np.abs(ref_logits)))
        ref_loss, ref_g = mlp_ref.loss_and_gradients(rows, labels, starts, w1, b1, w2, b2)
        g, loss = LN.gradients(LN.Dataset(rows, labels, starts))
        assert abs(loss - ref_loss) <= 1e-4 * max(1.0, abs(ref_loss))
        assert np.linalg.norm(g - ref_g) <= 2e-3 * np.linalg.norm(ref_g)
        tail = slice(128 * 256, None)
        assert np.linalg.norm(g[tail] - ref_g[tail]) <= 1e-3 * np.linalg.norm(ref_g[tail])


def _plain_witnesses(slot, cases, rate, D):
    res = T.search_batch(slot, [c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases], rate, D,
                         0, flags=T.FLAG_DETERMINISTIC)
    return res, [res.moves[c] if res.found[c] else [] for c in range(len(cases))]


@pytest.mark.gpu
def test_training_lowers_loss_and_ranks_taken_moves(slots):
    wl = W.workload_four_3cycles_n24(False)
    cases = wl.cases[:1024]
    _, wit = _plain_witnesses(slots[0], cases, wl.min_rate, wl.max_depth)
    ds = LN.build_dataset(cases, wl.min_rate, wl.max_depth, wit)
    assert ds.groups > 400
    LN.init(5)
    before = LN.group_accuracy(ds, LN.forward(ds.rows))
    losses = LN.train(ds, epochs=12, lr=2e-3, groups_per_batch=64)
    after = LN.group_accuracy(ds, LN.forward(ds.rows))
    chance = np.mean([1.0 / (ds.group_start[g + 1] - ds.group_start[g]) for g in range(ds.groups)])
    print(f"groups {ds.groups} rows {len(ds.rows)} losses {np.round(losses, 4)} accuracy {before:.3f} -> "
          f"{after:.3f} (chance {chance:.3f})", flush=True)
    # the label is the one sequence the plain search returned; other moves
    # often lead to a sequence too, so accuracy is bounded well below 1
    assert losses[-1] < 0.9 * losses[0]
    assert after >= before + 0.1 and after > 1.5 * chance


@pytest.mark.gpu
def test_learned_ordering_keeps_verdicts(slots):
    """Same verdicts as the oracle (C1 batch) and as the plain search (C2
    unoriented), every witness replays; deterministic mode: expansions to
    verdict of the learned ordering against the plain one."""
    wl2 = W.workload_four_3cycles_n24(False)
    train_cases = wl2.cases[:1024]
    _, wit = _plain_witnesses(slots[0], train_cases, wl2.min_rate, wl2.max_depth)
    LN.init(5)
    LN.train(LN.build_dataset(train_cases, wl2.min_rate, wl2.max_depth, wit), epochs=12, lr=2e-3,
             groups_per_batch=64)

    wl = W.workload_random_n12()
    want = oracle.search_many([c.pi for c in wl.cases], [c.spi for c in wl.cases], [c.n for c in wl.cases],
                              [c.even_cycles for c in wl.cases], [wl.min_rate] * len(wl.cases),
                              [wl.max_depth] * len(wl.cases), [0] * len(wl.cases), True, 0, 8)[0]
    for flags in (T.FLAG_DETERMINISTIC | _lib.FLAG_LEARNED, _lib.FLAG_LEARNED):
        got = T.search_batch(slots[0], [c.pi for c in wl.cases], [c.spi for c in wl.cases],
                             [c.even_cycles for c in wl.cases], wl.min_rate, wl.max_depth, 0, flags=flags)
        assert [bool(f) for f in got.found] == [w.found for w in want]
        for c, case in enumerate(wl.cases):
            if got.found[c]:
                assert_witness(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth, got.moves[c])

    test_cases = wl2.cases[1024:2048]
    plain, _ = _plain_witnesses(slots[0], test_cases, wl2.min_rate, wl2.max_depth)
    lrn = T.search_batch(slots[0], [c.pi for c in test_cases], [c.spi for c in test_cases],
                         [c.even_cycles for c in test_cases], wl2.min_rate, wl2.max_depth, 0,
                         flags=T.FLAG_DETERMINISTIC | _lib.FLAG_LEARNED)
    assert np.array_equal(plain.found, lrn.found)
    sat = plain.found
    pe = plain.counters["expansions"][sat].astype(np.float64)
    le = lrn.counters["expansions"][sat].astype(np.float64)
    print(f"satisfiable {int(sat.sum())}: expansions to verdict plain {pe.sum():.0f} learned {le.sum():.0f} "
          f"(ratio {pe.sum() / max(le.sum(), 1):.2f})", flush=True)


@pytest.mark.gpu
def test_graph_switches_between_plain_and_learned_kernels(slots):
    """One slot alternates graph rounds of the plain and the learned search
    kernel, with the number of rounds (and so the reset grid) changing between
    calls: each graph call equals its eager twin (deterministic mode: verdict,
    sequence and counters)."""
    from tdp1375_amd import train_ordering
    train_ordering.load(os.path.join(os.path.dirname(LN.__file__), "data", "ordering_mlp_d11.npz"))
    wl = W.workload_random_n12()
    det = T.FLAG_DETERMINISTIC
    G = _lib.FLAG_GRAPH
    plan = [(0, 96), (_lib.FLAG_LEARNED, 96), (0, 40), (_lib.FLAG_LEARNED, 200), (_lib.FLAG_LEARNED, 40), (0, 200)]
    for extra, count in plan:
        cases = wl.cases[:count]
        args = ([c.pi for c in cases], [c.spi for c in cases], [c.even_cycles for c in cases], wl.min_rate,
                wl.max_depth, 0)
        got = T.search_batch(slots[1], *args, flags=det | extra | G)
        eager = T.search_batch(slots[1], *args, flags=det | extra)
        assert np.array_equal(got.found, eager.found), (extra, count)
        assert np.array_equal(got.depth, eager.depth), (extra, count)
        assert [list(map(tuple, m)) for m in got.moves] == [list(map(tuple, m)) for m in eager.moves]
        for name in got.counters:
            assert np.array_equal(got.counters[name], eager.counters[name]), (extra, count, name)
