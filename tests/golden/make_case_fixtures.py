"""Writes tests/golden/cases.json from the reference's proof-output data files.

The reference ships its proof outputs as text lines ``spi->[(a b c), ...]``
(src/main/resources/cases/cases-comb.txt, cases-3,2.txt,
cases-oriented-7cycle.txt).  This script reads those data files as text and
stores, per line, the configuration's cycles and the witness symbol triples.
It also records each file's SHA-256, size and whether it ends with a
newline (case_files_meta.json), so that the case-file writer can be checked
byte for byte.  No reference code is run.  Run from the repository root:

    python tests/golden/make_case_fixtures.py /root/reference
"""
import hashlib
import json
import os
import re
import sys

FILES = {
    "comb": "src/main/resources/cases/cases-comb.txt",
    "3_2": "src/main/resources/cases/cases-3,2.txt",
    "oriented_7cycle": "src/main/resources/cases/cases-oriented-7cycle.txt",
}

_CYCLE = re.compile(r"\(([^()]*)\)")


def parse_cycles(text):
    return [[int(x) for x in re.split(r"[ ,]+", c.strip()) if x] for c in _CYCLE.findall(text)]


def main(ref_root):
    out = {}
    files = {}
    for key, rel in FILES.items():
        raw = open(os.path.join(ref_root, rel), "rb").read()
        files[key] = {"sha256": hashlib.sha256(raw).hexdigest(), "bytes": len(raw),
                      "trailing_newline": raw.endswith(b"\n")}
        entries = []
        with open(os.path.join(ref_root, rel)) as fh:
            for line in fh:
                line = line.strip()
                if not line:
                    continue
                lhs, rhs = line.split("->")
                if key == "oriented_7cycle":
                    # Silvaetal.java:151-165: the left side is pi (as an
                    # n-cycle) and spi is fixed to (0,3,4,1,5,2,6).
                    entries.append({"spi": [[0, 3, 4, 1, 5, 2, 6]],
                                    "pi": parse_cycles(lhs)[0],
                                    "witness": parse_cycles(rhs)})
                else:
                    entries.append({"spi": parse_cycles(lhs), "witness": parse_cycles(rhs)})
        out[key] = entries
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cases.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    # byte-level facts of the files, for the case-file writer's round trip
    with open(os.path.join(os.path.dirname(dst), "case_files_meta.json"), "w") as fh:
        json.dump(files, fh, indent=1)
    print({k: len(v) for k, v in out.items()}, "->", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
