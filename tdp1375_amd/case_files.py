"""Case files: the proof's output format ``spi->[(a b c), (d e f), ...]``.

One line per configuration: its cycles written as ``(a b c)`` one after the
other, ``->``, and the sorting sequence as symbol triples in brackets,
separated by ``", "``.  This is what AbstractSbtAlgorithm.loadSortings reads
(AbstractSbtAlgorithm.java:165-179: the line is trimmed and split at "->",
spaces in the left side become commas for MulticyclePermutation, the right
side loses its brackets and is split into cycles), and what the shipped
cases-*.txt hold, so a file written here from GPU search results can replace
or extend them.  For cases-oriented-7cycle.txt the left side is pi instead of
spi (Silvaetal.java:151-165); the format is the same.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass
from typing import Iterable, List, Sequence, Tuple

Cycles = List[List[int]]

_CYCLE = re.compile(r"\(([^()]*)\)")


@dataclass
class CaseLine:
    left: Cycles            # spi (or pi for the oriented 7-cycle file)
    sorting: Cycles         # the sequence of 3-cycles, in order


def _cycle(c: Sequence[int]) -> str:
    return "(" + " ".join(str(int(s)) for s in c) + ")"


def format_line(left: Sequence[Sequence[int]], sorting: Sequence[Sequence[int]]) -> str:
    return "".join(_cycle(c) for c in left) + "->[" + ", ".join(_cycle(m) for m in sorting) + "]"


def parse_line(line: str) -> CaseLine:
    lhs, rhs = line.strip().split("->")
    rhs = rhs.strip()
    if not (rhs.startswith("[") and rhs.endswith("]")):
        raise ValueError(f"malformed case line: {line!r}")
    parse = lambda t: [[int(x) for x in re.split(r"[ ,]+", c.strip()) if x] for c in _CYCLE.findall(t)]
    return CaseLine(parse(lhs), parse(rhs[1:-1]))


def dumps(lines: Iterable[CaseLine], trailing_newline: bool = True) -> str:
    text = "\n".join(format_line(l.left, l.sorting) for l in lines)
    return text + "\n" if trailing_newline and text else text


def loads(text: str) -> List[CaseLine]:
    return [parse_line(l) for l in text.splitlines() if l.strip()]


def write_case_file(path: str, lines: Iterable[CaseLine], trailing_newline: bool = True) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w", newline="\n") as fh:
        fh.write(dumps(lines, trailing_newline))
    os.replace(tmp, path)


def read_case_file(path: str) -> List[CaseLine]:
    with open(path) as fh:
        return loads(fh.read())


def from_results(configs: Sequence[Cycles], moves: Sequence) -> List[CaseLine]:
    """Lines for the configurations a search found a sequence for (moves[c]
    a list of symbol triples, or None when not found)."""
    return [CaseLine([list(map(int, c)) for c in cfg], [list(map(int, m)) for m in mv])
            for cfg, mv in zip(configs, moves) if mv]
