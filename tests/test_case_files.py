"""Case files ``spi->[(a b c), ...]`` (the proof's output format, read by
AbstractSbtAlgorithm.loadSortings, AbstractSbtAlgorithm.java:165-179).

CPU: the writer reproduces the shipped cases-comb.txt, cases-3,2.txt and
cases-oriented-7cycle.txt byte for byte (SHA-256 and size recorded from the
files by tests/golden/make_case_fixtures.py) from their parsed lines, and
reading what it writes gives the lines back.  GPU: the batched depth ladder's
results for the cases-3,2 configurations, written as a case file and read
back, replay to accepted states.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from tdp1375_amd import case_files as CF
from tdp1375_amd import proof_driver as PD
from helpers import MIN_RATE_15, assert_witness

HERE = os.path.dirname(__file__)


def load(name):
    with open(os.path.join(HERE, "golden", name)) as fh:
        return json.load(fh)


def test_writer_reproduces_the_shipped_files_byte_for_byte():
    cases, meta = load("cases.json"), load("case_files_meta.json")
    for key, entries in cases.items():
        lines = [CF.CaseLine([e["pi"]] if "pi" in e else e["spi"], e["witness"]) for e in entries]
        raw = CF.dumps(lines, meta[key]["trailing_newline"]).encode()
        assert len(raw) == meta[key]["bytes"], key
        assert hashlib.sha256(raw).hexdigest() == meta[key]["sha256"], key
        back = CF.loads(raw.decode())
        assert [(l.left, l.sorting) for l in back] == [(l.left, l.sorting) for l in lines], key


def test_parse_tolerates_spacing_and_rejects_malformed_lines(tmp_path):
    l = CF.parse_line("  (0 4 2)(1,5,3)->[(0 2 4), (1 5 3)]  ")
    assert l.left == [[0, 4, 2], [1, 5, 3]] and l.sorting == [[0, 2, 4], [1, 5, 3]]
    with pytest.raises(ValueError):
        CF.parse_line("(0 4 2)->(0 2 4)")
    p = tmp_path / "c.txt"
    CF.write_case_file(str(p), [l, l])
    assert p.read_text() == "(0 4 2)(1 5 3)->[(0 2 4), (1 5 3)]\n" * 2


@pytest.mark.gpu
def test_ladder_results_written_as_a_case_file_replay(slots, tmp_path):
    configs = [e["spi"] for e in load("cases.json")["3_2"]]
    res = PD.sort_batch(slots[0], configs, 1.5, 11)
    assert res.found.all()
    path = str(tmp_path / "cases-3,2.gpu.txt")
    assert PD.write_cases(path, configs, res) == len(configs)
    back = CF.read_case_file(path)
    assert [l.left for l in back] == configs
    for l in back:
        spi = PD.one_line(l.left)
        pi = list(range(len(spi)))
        # symbol triples -> positions on the replayed pi (GPUSortingSearch.java:65-87 in reverse)
        positions, cur = [], list(pi)
        for a, b, c in l.sorting:
            i, j, k = sorted((cur.index(a), cur.index(b), cur.index(c)))
            positions.append((i, j, k))
            cur = cur[:i] + cur[j:k] + cur[i:j] + cur[k:]
        assert_witness(pi, spi, PD.odd_cycles(l.left), MIN_RATE_15, 11, positions)
