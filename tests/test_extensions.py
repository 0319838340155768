"""Extension generators (Extensions.java:95-203, Combinations.java:64-94) and
the level-by-level SortOrExtend exploration over them.

The reference has no tests of its own for these; they are pinned by
* the configurations its proof pages name (index.html:44-59, fixture written
  by tests/golden/make_extension_fixtures.py): the canonical spi of every
  DFS root and bad small component, and the full configuration
  (0 9 2)(1 11 4)(3 8 6)(5 10 7) that the DFS reaches and that does not allow
  an 11/8-sequence;
* closed forms: |type 1| = |open gates| * (n(n+1)/2 - 1), |type 2| =
  C(n+2, 3) - n for a full configuration and 0 otherwise, |combinations| =
  6n;
* structure: the inserted symbols sit where the info string says, and
  deleting them gives back the parent; a grown cycle stays unoriented with at
  most two open gates after;
* the comment of Extensions.java:192: a 3-cycle turned into an oriented
  5-cycle is kept exactly when that 5-cycle admits no 2-move (brute force
  over every transposition).
GPU: the exploration of the DFS tree to three levels, every canonical form
searched by the batched ladder, every sorting replayed, the roots and the
documented full configuration without a sorting.
"""
import itertools
import json
import os
import random
from math import comb

import numpy as np
import pytest

from oracle import canon_ref as C
from oracle import extensions_ref as R
from tdp1375_amd import extensions as E
from tdp1375_amd import proof_driver as PD

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "extension_roots.json")


def fixture():
    with open(GOLDEN) as fh:
        return json.load(fh)


def canonical_string(cycles):
    n = E.size(cycles)
    sig, _, _ = C.canonical(cycles, list(range(n)))
    return PD.spi_string(C.of_signature(sig))


def perm(cycles, n):
    p = list(range(n))
    for c in cycles:
        for t, s in enumerate(c):
            p[s] = c[(t + 1) % len(c)]
    return p


def odd_cycles(p):
    seen, k = [False] * len(p), 0
    for s in range(len(p)):
        if not seen[s]:
            L, x = 0, s
            while not seen[x]:
                seen[x], x, L = True, p[x], L + 1
            k += L % 2
    return k


def restrict(cycles, symbols):
    """The configuration on the given symbols (every cycle mapping inside
    them), relabelled 0..m-1 in order."""
    keep = sorted(symbols)
    idx = {s: i for i, s in enumerate(keep)}
    return perm([[idx[s] for s in c] for c in cycles if s_in(c, idx)], len(keep))


def s_in(c, idx):
    inside = [s in idx for s in c]
    assert all(inside) or not any(inside), "a cycle straddles the inserted and the parent symbols"
    return all(inside)


def has_2move(p):
    """A transposition (a b c), a < b < c, raising the odd-cycle count by 2
    (spi' = spi . (a b c)^-1)."""
    o = odd_cycles(p)
    for a, b, c in itertools.combinations(range(len(p)), 3):
        ti = {a: c, c: b, b: a}
        if odd_cycles([p[ti.get(x, x)] for x in range(len(p))]) == o + 2:
            return True
    return False


def random_configs(seed, count, lo=3, hi=12):
    rng = random.Random(seed)
    out = []
    while len(out) < count:
        n = rng.randint(lo, hi)
        p = list(range(n))
        rng.shuffle(p)
        if all(p[x] != x for x in range(n)):
            out.append(C.cycles_of(p))
    return out


# ---------------------------------------------------------------- CPU
def test_roots_and_their_canonical_forms_match_the_proof_pages():
    g = fixture()
    assert R.DFS_ROOTS == E.DFS_ROOTS and R.BAD_SMALL_COMPONENTS == E.BAD_SMALL_COMPONENTS
    assert [PD.spi_string(E.parse(s)) for s in g["dfs_roots"]] == [PD.spi_string(r) for r in E.DFS_ROOTS]
    assert [PD.spi_string(E.parse(s)) for s in g["bad_small_components"]] == \
        [PD.spi_string(c) for c, _ in E.BAD_SMALL_COMPONENTS]
    names = {"the bad oriented $5$-cycle": E.ORIENTED_5_CYCLE,
             "the unoriented interleaving pair": E.INTERLEAVING_PAIR,
             "the unoriented intersecting pair": E.INTERSECTING_PAIR,
             "the unoriented necklace of size $4$": E.NECKLACE_SIZE_4,
             "the twisted necklace of size $4$": E.TWISTED_NECKLACE_SIZE_4,
             "the unoriented necklace of size $5$": E.NECKLACE_SIZE_5,
             "the unoriented necklace of size $6$": E.NECKLACE_SIZE_6}
    for name, cfg in names.items():
        assert canonical_string(cfg) == g["dfs_canonical_by_name"][name], name
        if name in g["comb_spi_by_name"]:
            assert PD.spi_string(cfg) == g["comb_spi_by_name"][name], name


def test_signature_matches_the_oracle_and_round_trips():
    for cyc in random_configs(1, 200, 3, 24):
        n = E.size(cyc)
        for M in (E, R):
            sig = M.signature(cyc)
            assert np.array_equal(sig, C.signature(cyc, list(range(n))))
            assert perm(M.of_signature(sig), n) == perm(cyc, n)
            assert M.of_signature(sig) == R.of_signature(sig)


def test_generators_equal_the_oracle():
    """The batched generators of the package against the one-at-a-time
    restatement: the same (info, configuration) lists in the same order, on
    random configurations, the roots and a sample of the first two levels of
    the DFS tree (oriented cycles, full configurations, n up to 15)."""
    cfgs = list(R.DFS_ROOTS) + [c for c, _ in R.BAD_SMALL_COMPONENTS] + random_configs(5, 40, 3, 14)
    level = list(R.DFS_ROOTS)
    for depth in range(2):
        level = [x for c in level for _, x in R.dfs_extensions(c)][::11 if depth == 0 else 97]
        cfgs += level
    for c in cfgs:
        assert E.open_gates(c) == R.open_gates(c), c
        for name in ("type1_extensions", "type2_extensions", "type3_extensions", "combination_extensions"):
            assert getattr(E, name)(c) == getattr(R, name)(c), (name, c)
        # the one-line rows explore works on: the same configurations in the same order
        for api, rows_of in ((R.dfs_extensions, E.dfs_extension_rows),
                             (R.combination_extensions, E.combination_extension_rows)):
            rows, ns = rows_of(E.one_line_of(c))
            assert [list(map(int, rows[i, :ns[i]])) for i in range(len(ns))] == \
                [perm(x, E.size(x)) for _, x in api(c)], (api.__name__, c)


def test_extension_counts_follow_the_closed_forms():
    for cyc in random_configs(2, 40) + list(E.DFS_ROOTS):
        n = E.size(cyc)
        gates = R.open_gates(cyc)
        assert len(R.type1_extensions(cyc)) == len(gates) * (n * (n + 1) // 2 - 1)
        assert len(R.type2_extensions(cyc)) == (0 if gates else comb(n + 2, 3) - n)
        assert len(R.combination_extensions(cyc)) == 6 * n


def test_inserted_symbols_sit_where_the_info_says_and_removing_them_gives_the_parent():
    for cyc in random_configs(3, 12, 3, 9) + list(E.DFS_ROOTS):
        n = E.size(cyc)
        parent = perm(cyc, n)
        for gen in (R.type1_extensions, R.type2_extensions):
            for info, ext in gen(cyc):
                a, b, c = sorted(int(x.split("=")[1]) for x in info.split())
                new = {a, b + 1, c + 2}
                assert sorted(ext[-1]) == sorted(new), info         # the new cycle has the last label
                assert len(ext) == len(cyc) + 1
                assert restrict(ext, set(range(n + 3)) - new) == parent, info
                assert restrict(ext, new) in ([1, 2, 0], [2, 0, 1]) and not has_2move(restrict(ext, new))
        for info, ext in R.type3_extensions(cyc):
            a, b = (int(x.split("=")[1]) for x in info.split(",")[0].split())
            grown = [c for c in ext if a in c and b + 1 in c]
            assert len(grown) == 1 and len(ext) == len(cyc), info
            rest = [c for c in ext if c is not grown[0]]
            old = sorted(set(grown[0]) - {a, b + 1})
            # dropping the two new symbols of the grown cycle gives the parent
            shrunk = [[s for s in c if s not in (a, b + 1)] for c in ext]
            keep = sorted(set(range(n + 2)) - {a, b + 1})
            assert restrict(shrunk, keep) == parent, info
            if "turn oriented" not in info:
                sig = R.signature(ext)
                lab = sig[grown[0][0]]
                assert all(sig[s] == lab for s in grown[0]) and float(lab) % 1 == 0, info   # unoriented
                assert len(R.open_gates(ext)) <= 2, info
            else:
                assert len(old) == 3 and len(grown[0]) == 5, info


def test_turned_oriented_5_cycles_are_kept_exactly_when_they_admit_no_2_move():
    """Extensions.java:192: otherwise the configuration accepts a 2-move."""
    kept = rejected = 0
    for cyc in list(E.DFS_ROOTS) + [c for c, _ in E.BAD_SMALL_COMPONENTS[:4]] + random_configs(4, 10, 6, 12):
        sig = R.signature(cyc)
        n = len(sig)
        by_label = {}
        for i in range(n):
            by_label.setdefault(int(sig[i]), []).append(i)
        cyc_of = {s: c for c in cyc for s in c}
        for label, pos in by_label.items():
            if len(pos) != 3 or float(sig[pos[0]]) % 1:
                continue
            fr = np.asarray([np.float32(np.float32(f) + np.float32(label)) for f in R._RANK_ORDER], np.float32)
            for a in range(n):
                for b in range(a, n):
                    if R._remains_unoriented(pos, a, b):
                        continue
                    prime = R._turn_oriented(cyc_of[pos[0]], label, sig, a, b)
                    five = [c for c in R.of_signature(prime) if len(c) == 5 and a in c][0]
                    alone = restrict([five], set(five))
                    ok = R._in_rank_order(prime, fr)
                    assert ok == (not has_2move(alone)), (cyc, label, a, b)
                    kept += ok
                    rejected += not ok
    assert kept and rejected


def test_documented_full_configuration_is_two_levels_below_the_roots():
    """index.html:59: the DFS reached (0 9 2)(1 11 4)(3 8 6)(5 10 7)."""
    g = fixture()
    target = C.config_hash(E.parse(g["full_configuration_without_11_8"]), list(range(12)))
    level = list(E.DFS_ROOTS)
    for _ in range(2):
        nxt = {}
        for c in level:
            for _, x in R.dfs_extensions(c):
                nxt.setdefault(tuple(perm(x, E.size(x))), x)
        level = list(nxt.values())
    hashes = {C.config_hash(c, list(range(12))) for c in level if E.size(c) == 12 and R.is_full(c)}
    assert target in hashes


def test_exploration_searches_each_canonical_form_once_and_extends_only_bad_ones(monkeypatch):
    """Host logic of explore with the oracle's canonical form and a stand-in
    verdict (a sorting exists iff the configuration has an oriented cycle)."""
    from tdp1375_amd import canonical as K
    searched = []

    def canon(rows, ns):
        out = np.zeros((len(ns), 64), np.uint8)
        for i, n in enumerate(ns):
            cyc = C.cycles_of([int(x) for x in rows[i, :n]])
            sig, _, _ = C.canonical(cyc, list(range(n)))
            out[i, :n] = perm(C.of_signature(sig), n)
        return None, None, None, out

    def sort_rows(slot, rows, ns, ratio, depth, e0=None, norms=None, flags=0):
        configs = [C.cycles_of([int(x) for x in rows[i, :ns[i]]]) for i in range(len(ns))]
        searched.extend(tuple(perm(c, E.size(c))) for c in configs)
        found = np.array([any(float(v) % 1 for v in E.signature(c)) for c in configs])
        return PD.LadderResult(found, np.full(len(configs), 4), [[(0, 1, 2)] if f else None for f in found], 1)

    monkeypatch.setattr(K, "canonical_rows", canon)
    monkeypatch.setattr(PD, "sort_rows", sort_rows)
    res = E.explore([E.INTERLEAVING_PAIR, E.INTERLEAVING_PAIR], E.dfs_extension_rows, 2)
    assert len(searched) == len(set(searched)) == len(res.verdicts)
    assert res.levels[0].configurations == 1 and res.levels[0].extended == 1
    lv = res.levels[1]
    kids = {tuple(perm(x, E.size(x))) for _, x in E.dfs_extensions(E.INTERLEAVING_PAIR)}
    assert lv.configurations == len(kids)
    assert lv.extended == sum(1 for k in kids if not any(float(v) % 1 for v in E.signature(C.cycles_of(list(k)))))
    res = E.explore([E.INTERLEAVING_PAIR], E.dfs_extension_rows, 3, max_configurations=10)
    assert res.truncated and len(res.levels) == 1


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_dfs_tree_three_levels_on_the_gpu(slots):
    from helpers import assert_witness
    g = fixture()
    res = E.explore(E.DFS_ROOTS, E.dfs_extension_rows, 3, 1.375, 11, slot=slots[0])
    assert [l.configurations for l in res.levels][:2] == [3, 123]
    min_rate = float(np.float32(2.0) / np.float32(1.375))
    for r in E.DFS_ROOTS:                      # the roots do not allow 11/8 (index.html:52-56)
        k = tuple(perm(C.of_signature(C.canonical(r, list(range(E.size(r))))[0]), E.size(r)))
        assert res.verdicts[k] is False
    full = E.parse(g["full_configuration_without_11_8"])
    assert res.verdicts[tuple(perm(full, 12))] is False
    for k, moves in res.sortings.items():
        cyc = C.cycles_of(list(k))
        pi = list(range(len(k)))
        positions, cur = [], list(pi)
        for a, b, c in moves:
            i, j, l = sorted((cur.index(a), cur.index(b), cur.index(c)))
            positions.append((i, j, l))
            cur = cur[:i] + cur[j:l] + cur[i:j] + cur[l:]
        assert_witness(pi, list(k), PD.odd_cycles(cyc), min_rate, 11, positions)


@pytest.mark.gpu
def test_dfs_tree_extended_on_worker_processes_equals_one_process(slots):
    # levels extended on four host processes: the same levels, canonical
    # forms and verdicts as one process (chunks joined in order)
    one = E.explore(E.DFS_ROOTS, E.dfs_extension_rows, 4, 1.375, 11, slot=slots[0])
    four = E.explore(E.DFS_ROOTS, E.dfs_extension_rows, 4, 1.375, 11, slot=slots[0], workers=4)
    assert [vars(l)["configurations"] for l in one.levels] == [vars(l)["configurations"] for l in four.levels]
    assert [l.new_canonical for l in one.levels] == [l.new_canonical for l in four.levels]
    assert one.verdicts == four.verdicts


@pytest.mark.gpu
def test_combination_tree_runs_to_completion_on_the_gpu(slots):
    """index.html:63: no combination of bad small components with 3-norm
    greater than 7 was extended, and the search over combinations ends."""
    res = E.explore([c for c, _ in E.BAD_SMALL_COMPONENTS], E.combination_extension_rows, 8, 1.375, 11,
                    slot=slots[0])
    assert not res.truncated
    assert res.levels[-1].extended == 0                  # the last level has a sorting for everything
    assert max(l.max_bad_three_norm for l in res.levels) == 7
    assert res.levels[0].extended == 6                   # every bad small component lacks a sorting


@pytest.mark.gpu
def test_dfs_verdicts_equal_the_oracle_ladder(slots):
    """Every canonical form of the first four DFS levels (6,249) gets the same
    verdict from the batched GPU ladder as from the oracle's search run
    through the same depth ladder (SortOrExtend.java:89-105) on the CPU."""
    import oracle
    res = E.explore(E.DFS_ROOTS, E.dfs_extension_rows, 4, 1.375, 11, slot=slots[0])
    keys = list(res.verdicts)
    assert len(keys) == 3 + 20 + 535 + 5691
    ns = np.array([len(k) for k in keys], np.int32)
    rows = np.zeros((len(keys), 64), np.uint8)
    for c, k in enumerate(keys):
        rows[c, :len(k)] = k
    e0 = PD.odd_cycle_counts(rows, ns)
    min_rate = float(np.float32(2) / np.float32(1.375))
    ub = np.floor(((ns - e0) // 2).astype(np.float32) * (np.float32(2) / np.float32(min_rate)))
    ladders = [PD.depth_ladder(float(u), 11) for u in ub]
    found = np.zeros(len(keys), bool)
    ident = np.broadcast_to(np.arange(64, dtype=np.uint8), (len(keys), 64))
    for rung in range(3):
        todo = [c for c in range(len(keys)) if not found[c] and rung < len(ladders[c])]
        if not todo:
            break
        out, _ = oracle.search_many(ident[todo], rows[todo], ns[todo], e0[todo], [min_rate] * len(todo),
                                    [ladders[c][rung] for c in todo], [0] * len(todo), True, 0, 8)
        for q, c in enumerate(todo):
            found[c] |= out[q].found
    assert [res.verdicts[k] for k in keys] == found.tolist()
