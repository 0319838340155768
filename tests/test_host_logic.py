"""CPU tests of the library's host side through the C ABI (no GPU compute)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import _lib
from tdp1375_amd import workloads as W
from helpers import MIN_RATE_15, MIN_RATE_1375

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "tdp1375.h")) as fh:
        return re.findall(r"^TDP_API [^(]*?\b(tdp_\w+)\(", fh.read(), flags=re.M)


def test_library_exports_every_declared_symbol():
    lib = T.load()
    declared = header_symbols()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(lib, name), name
    for name in _lib.EXPORTED_SYMBOLS:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\S+)", out))
    assert set(declared) <= exported
    assert {s for s in _lib.EXPORTED_SYMBOLS if s.startswith("Java_")} <= exported


def test_host_parity_delta_equals_recount():
    lib = T.load()
    rng = np.random.default_rng(11)
    for it in range(80):
        n = int(rng.integers(3, 64))
        pi = [int(x) for x in rng.permutation(n)]
        spi = [int(x) for x in rng.permutation(n)] if it % 2 else W.one_line(
            [list(map(int, rng.permutation(n)[: int(rng.integers(2, 8))]))], n)
        want = oracle.children_even(pi, spi)
        got = np.zeros(len(want), np.int32)
        _lib.check(lib.tdp_host_eval_children(_lib.u8(pi, 64), _lib.u8(spi, 64), n,
                                              got.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        assert np.array_equal(got, want)


@pytest.mark.parametrize("rate", [MIN_RATE_1375, MIN_RATE_15, float(np.float32(2) / np.float32(1.25)), 1.0])
def test_predicate_tables_equal_float32_predicates(rate):
    lib = T.load()
    acc = (ctypes.c_uint64 * 16)()
    rch = (ctypes.c_uint64 * 16)()
    for n in range(3, 64, 4):
        for e0 in range(0, n + 1, 5):
            for D in (1, 4, 9, 11, 15):
                for full in (0, 1):
                    _lib.check(lib.tdp_host_predicate_tables(n, e0, rate, D, full, acc, rch))
                    for d in range(16):
                        for e in range(n + 1):
                            assert bool(acc[d] >> e & 1) == oracle.satisfies_rate(e, d, n, e0, rate, full)
                            assert bool(rch[d] >> e & 1) == oracle.can_still_reach(e, d, n, e0, rate, D, full)


def test_argument_errors_without_gpu():
    """Validation happens before any device work, with the reference's messages."""
    with pytest.raises(T.TdpError, match="performGpuSortingSearch: n exceeds MAX_N"):
        T.search_for_sorting_seq(list(range(64)), list(range(64)), 0, 1.0, 4, False, 0)
    with pytest.raises(T.TdpError, match="performGpuSortingSearch: invalid maximumDepth"):
        T.search_for_sorting_seq(list(range(8)), list(range(8)), 8, 1.0, 16, False, 0)
    with pytest.raises(T.TdpError, match="invalid slot"):
        T.search_for_sorting_seq(list(range(8)), list(range(8)), 8, 1.0, 4, False, 123)
    with pytest.raises(T.TdpError, match="n must be at least 1"):
        T.filter_pi_permutations([0], 1, [0], 1, 0, False)
    with pytest.raises(T.TdpError, match="n exceeds PI_FILTER_MAX_N"):
        T.filter_pi_permutations([0], 1, [0], 1, 64, False)
    for bad in (-1, 17):
        with pytest.raises(T.TdpError, match="batch must be in 0..16"):
            T.set_expansion_batch(bad)
    T.cancel_sort(-5)   # never throws (slot_manager.cu:100-125)


def test_positions_to_symbols_replay():
    """GPUSortingSearch.java:65-87: positions are replayed on a copy of pi."""
    pi = [0, 1, 2, 3, 4, 5]
    moves = T.positions_to_symbols(pi, [(0, 2, 4), (1, 3, 5)])
    assert moves[0] == (0, 2, 4)
    # after the first block swap pi = [2, 3, 0, 1, 4, 5]
    assert moves[1] == (3, 1, 5)
    assert pi == [0, 1, 2, 3, 4, 5]


def test_plus2_children_stay_in_one_cycle_or_two_even_cycles():
    """Premise of the kernel's cycle-local enumeration: a child with two more
    odd cycles moves three symbols of one oriented cycle (two or more ascents
    of the pi-positions along spi), or two symbols of one even cycle and one
    of another even cycle (checked with the oracle's O(n) recount of every
    child, on nodes with and without even cycles)."""
    rng = np.random.default_rng(5)
    for it in range(80):
        n = int(rng.integers(5, 40))
        lengths, left = [], n
        odd_only = it % 2 == 0
        while left > 0:
            L = int(rng.choice([1, 3, 5, 7] if odd_only else [1, 2, 3, 4, 5, 6]))
            L = min(L, left)
            if odd_only and L % 2 == 0:
                L -= 1
            lengths.append(L)
            left -= L
        sym = [int(x) for x in rng.permutation(n)]
        cycles, p = [], 0
        for L in lengths:
            cycles.append(sym[p:p + L])
            p += L
        spi = W.one_line(cycles, n)
        pi = [int(x) for x in rng.permutation(n)]
        e = W.odd_cycle_count(spi)
        cyc_of = {s: c for c, cyc in enumerate(cycles) for s in cyc}
        even = {c for c, cyc in enumerate(cycles) if len(cyc) % 2 == 0}
        E = sum(len(cycles[c]) for c in even)
        pos = {x: q for q, x in enumerate(pi)}
        ascents = [sum(1 for t in range(len(c)) if pos[c[(t + 1) % len(c)]] > pos[c[t]]) for c in cycles]
        kids = oracle.children_even(pi, spi)
        t = 0
        plus2 = 0
        for i in range(n):
            for j in range(i + 1, n):
                for k in range(j + 1, n):
                    if kids[t] == e + 2:
                        plus2 += 1
                        cs = sorted({cyc_of[pi[i]], cyc_of[pi[j]], cyc_of[pi[k]]})
                        assert len(cs) == 1 or (len(cs) == 2 and set(cs) <= even), (it, i, j, k)
                        assert len(cs) == 2 or ascents[cs[0]] >= 2, (it, i, j, k)
                    t += 1
        bound = sum(len(c) * (len(c) - 1) * (len(c) - 2) // 6 for c, a in zip(cycles, ascents) if a >= 2)
        bound += sum(len(cycles[c]) * (len(cycles[c]) - 1) // 2 * (E - len(cycles[c])) for c in even)
        assert plus2 <= bound


def test_the_package_never_imports_the_oracle():
    """The oracle is the tests' checker only: no module of the package
    imports it, and importing the package (all modules) leaves it unloaded."""
    import ast
    import glob
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for path in glob.glob(os.path.join(root, "tdp1375_amd", "**", "*.py"), recursive=True):
        tree = ast.parse(open(path).read())
        for node in ast.walk(tree):
            names = [a.name for a in node.names] if isinstance(node, ast.Import) else \
                [node.module or ""] if isinstance(node, ast.ImportFrom) else []
            assert not any(n == "oracle" or n.startswith("oracle.") for n in names), path
    mods = [os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(root, "tdp1375_amd", "*.py"))
            if not p.endswith("__init__.py")]
    code = ("import sys, importlib\n" + "".join(f"importlib.import_module('tdp1375_amd.{m}')\n" for m in mods) +
            "print(any(k == 'oracle' or k.startswith('oracle.') for k in sys.modules))")
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "False"


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so,
    before loading the library or touching a device."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr and res.stdout == ""
