"""Seeded synthetic search inputs for the benchmark and parity workloads.

Every generator draws from SplitMix64 (Steele, Lea, Flood 2014) and shuffles
with Fisher-Yates, so the same seed gives the same cases on every machine.

A case is ``(pi, spi)``: ``pi`` is the linear arrangement (a cut of the n-cycle
pi) and ``spi`` the one-line sigma.pi^-1 the reference searches on
(GPUSortingSearch.java:47-75, SortOrExtend.java:90-94).  The proof generator
always searches with the identity ``pi`` (MulticyclePermutation.CANONICAL_PI_BYTE,
SortOrExtend.java:90).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, bound: int) -> int:
        return self.next() % bound

    def shuffle(self, items: List[int]) -> List[int]:
        for i in range(len(items) - 1, 0, -1):
            j = self.below(i + 1)
            items[i], items[j] = items[j], items[i]
        return items


def one_line(cycles: Sequence[Sequence[int]], n: int) -> List[int]:
    """One-line notation of a product of disjoint cycles (fixed points implicit)."""
    s = list(range(n))
    for c in cycles:
        for i, x in enumerate(c):
            s[x] = c[(i + 1) % len(c)]
    return s


def odd_cycle_count(spi: Sequence[int]) -> int:
    """Number of odd-length cycles (the reference's "even cycles")."""
    n = len(spi)
    seen = [False] * n
    count = 0
    for s in range(n):
        if seen[s]:
            continue
        size, cur = 0, s
        while not seen[cur]:
            seen[cur] = True
            size += 1
            cur = spi[cur]
        count += size & 1
    return count


def cycles_of(spi: Sequence[int]) -> List[List[int]]:
    n = len(spi)
    seen = [False] * n
    out = []
    for s in range(n):
        if seen[s]:
            continue
        c, cur = [], s
        while not seen[cur]:
            seen[cur] = True
            c.append(cur)
            cur = spi[cur]
        out.append(c)
    return out


def spi_from_pi(pi: Sequence[int]) -> List[int]:
    """spi = sigma . pi^-1 with sigma = (0 1 ... n-1): spi[pi_t] = pi_{t-1} + 1 mod n."""
    n = len(pi)
    spi = [0] * n
    for t in range(n):
        spi[pi[t]] = (pi[(t - 1) % n] + 1) % n
    return spi


@dataclass
class Case:
    pi: List[int]
    spi: List[int]
    label: str = ""

    @property
    def n(self) -> int:
        return len(self.pi)

    @property
    def even_cycles(self) -> int:
        return odd_cycle_count(self.spi)


@dataclass
class Workload:
    name: str
    cases: List[Case]
    max_ratio: float
    max_depth: int
    dedup_table_size: int
    queue_mb: int
    slots: int
    full_sorting: bool = False
    notes: dict = field(default_factory=dict)

    @property
    def min_rate(self) -> float:
        # SortOrExtend.java:43: minRate = 2 / maxRatio in Java float arithmetic
        return float(np.float32(2.0) / np.float32(self.max_ratio))


def random_permutations(count: int = 512, n: int = 12, seed: int = 0) -> List[Case]:
    """pi = [0] + shuffle(1..n-1); spi = sigma.pi^-1 (the 512 x n=12 workload)."""
    rng = SplitMix64(seed)
    cases = []
    for _ in range(count):
        pi = [0] + rng.shuffle(list(range(1, n)))
        cases.append(Case(pi, spi_from_pi(pi), "random"))
    return cases


def four_three_cycles(count: int = 2048, n: int = 24, seed: int = 1, oriented: bool = True) -> List[Case]:
    """Four 3-cycles on 12 of the n=24 symbols plus 12 fixed points, identity pi.

    As worded each 3-cycle is (a b c) with a<b<c, which is oriented with respect
    to the identity and therefore a 2-move; oriented=False writes (a c b).
    """
    rng = SplitMix64(seed)
    cases = []
    for _ in range(count):
        sym = rng.shuffle(list(range(n)))
        cycles = []
        for t in range(4):
            a, b, c = sorted(sym[3 * t: 3 * t + 3])
            cycles.append((a, b, c) if oriented else (a, c, b))
        cases.append(Case(list(range(n)), one_line(cycles, n), "oriented" if oriented else "unoriented"))
    return cases


def odd_cycle_mix(count: int = 4096, seed: int = 3, n_lo: int = 9, n_hi: int = 36,
                  oriented_fraction_pct: int = 30) -> List[Case]:
    """Uneven batch: n ~ U{n_lo..n_hi}, k ~ U{3..12} odd cycles of lengths in {3,5,7}.

    Lengths are redrawn while they do not fit n; after 16 failed draws k is
    decremented.  Symbols are shuffled, 3-cycles are written unoriented (a c b)
    except the first one in ``oriented_fraction_pct`` % of the cases; 5- and
    7-cycles keep the shuffled symbol order.  pi is the identity.
    """
    rng = SplitMix64(seed)
    cases = []
    for _ in range(count):
        n = n_lo + rng.below(n_hi - n_lo + 1)
        k = 3 + rng.below(10)
        tries = 0
        while True:
            lengths = [(3, 5, 7)[rng.below(3)] for _ in range(k)]
            if sum(lengths) <= n:
                break
            tries += 1
            if tries == 16:
                k -= 1
                tries = 0
        sym = rng.shuffle(list(range(n)))
        want_oriented = rng.below(100) < oriented_fraction_pct
        cycles, p, placed = [], 0, False
        for L in lengths:
            cyc = sym[p:p + L]
            p += L
            if L == 3:
                a, b, c = sorted(cyc)
                if want_oriented and not placed:
                    cyc = [a, b, c]
                    placed = True
                else:
                    cyc = [a, c, b]
            cycles.append(cyc)
        cases.append(Case(list(range(n)), one_line(cycles, n), "oriented3" if placed else "plain"))
    return cases


# The six "bad small components" of the 11/8 proof (Combinations.java:31-46):
# none of them admits an (11/8)-sequence.
BAD_SMALL_COMPONENTS = {
    "oriented_5_cycle": [(0, 3, 1, 4, 2)],
    "interleaving_pair": [(0, 4, 2), (1, 5, 3)],
    "necklace_4": [(0, 10, 2), (1, 5, 3), (4, 8, 6), (7, 11, 9)],
    "twisted_necklace_4": [(0, 7, 5), (1, 11, 9), (2, 6, 4), (3, 10, 8)],
    "necklace_5": [(0, 4, 2), (1, 14, 12), (3, 7, 5), (6, 10, 8), (9, 13, 11)],
    "necklace_6": [(0, 16, 2), (1, 5, 3), (4, 8, 6), (7, 11, 9), (10, 14, 12), (13, 17, 15)],
}


def bad_small_components() -> List[Case]:
    out = []
    for name, cyc in BAD_SMALL_COMPONENTS.items():
        n = 1 + max(max(c) for c in cyc)
        out.append(Case(list(range(n)), one_line(cyc, n), name))
    return out


def workload_random_n12() -> Workload:
    return Workload("random-n12", random_permutations(512, 12, 0), 1.375, 6, 1 << 20, 64, 1)


def workload_four_3cycles_n24(oriented: bool = True) -> Workload:
    return Workload("four-3cycles-n24" + ("" if oriented else "-unoriented"),
                    four_three_cycles(2048, 24, 1, oriented), 1.375, 9, 1 << 24, 3072, 2)


def workload_odd_cycle_mix_d11(count: int = 4096) -> Workload:
    return Workload("odd-cycle-mix-n9-36-d11", odd_cycle_mix(count, 3), 1.375, 11, 1 << 27, 16384, 1)


FILTER_ORIENTED_5_CYCLE = dict(cycle=[0, 3, 1, 4, 2], triple=[0, 1, 2])
