"""Shared test helpers: case-file fixtures and replay checks."""
import json
import os

import numpy as np

import oracle
from tdp1375_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MIN_RATE_1375 = float(np.float32(2.0) / np.float32(1.375))
MIN_RATE_15 = float(np.float32(2.0) / np.float32(1.5))


def load_cases():
    with open(os.path.join(GOLDEN, "cases.json")) as fh:
        return json.load(fh)


def case_inputs(entry):
    """(pi, spi) of a case-file entry (identity pi unless the file gives one)."""
    n = sum(len(c) for c in entry["spi"])
    spi = W.one_line(entry["spi"], n)
    if "pi" in entry:
        pi = list(entry["pi"])
    else:
        pi = list(range(n))
    return pi, spi


def replay_gain(pi, spi, positions):
    """Odd-cycle count after replaying position moves (-1 when invalid)."""
    e, _, _ = oracle.replay_positions(list(pi), list(spi), positions)
    return e


def assert_witness(pi, spi, e0, min_rate, max_depth, positions, full=False):
    """A returned sequence must replay to an accepted state within max_depth."""
    assert 1 <= len(positions) <= max_depth
    e = replay_gain(pi, spi, positions)
    assert e >= 0, "invalid move in witness"
    assert oracle.satisfies_rate(e, len(positions), len(pi), e0, min_rate, full), (
        f"witness does not satisfy the rate: gain {e - e0} in {len(positions)} moves")


def load_serving_pool():
    """(satisfiable cases, unsatisfiable cases, float32 min_rate, fixture) of
    the oracle-labelled serving pool (serving_pool_n9-48-d11.json)."""
    from tdp1375_amd import serving as S
    with open(os.path.join(GOLDEN, "serving_pool_n9-48-d11.json")) as fh:
        doc = json.load(fh)
    cases = S.pool_cases(doc["generated"], doc["seed"])
    sat = [cases[i] for i in doc["satisfiable"]]
    unsat = [cases[i] for i in doc["unsatisfiable"]]
    rate = float(np.float32(2.0) / np.float32(doc["max_ratio"]))
    return sat, unsat, rate, doc
