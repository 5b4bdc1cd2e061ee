"""Writes seg_cases.json: ff_seg's own stdout for edge-case ranges.

Runs the reference program built from reference/segmentedSieve.C by
oracle/Makefile (oracle/_ref/ff_seg and its PRIMEMEM variants) in a fresh
directory per case (ff_seg refuses to run where ./ffPlayground exists) and
stores the timing-normalised stdout and exit status.  Cases:
  * tiny ranges, where ff_seg's 31-thread split evaluates sumLimit+1 or skips
    a sum (segmentedSieve.C:706-761);
  * positional edge cases (one argument, none, even or small sumStart,
    sumStart > sumLimit; :844-855);
  * the capped geometry (:148-176) with PRIMEMEM 3000 B ("just sayin'"
    line, third header line, search split at pmapSumLimit :905-912) and
    PRIMEMEM 1000 B (the fatal message, exit status 255).

Run from the repo root after `make -C oracle`:  python tests/golden/make_seg_fixtures.py
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref")
NORM = re.compile(r"(Prime|Freudenthal) time: [0-9]+")

CASES = [
    # (binary, args, primemem)
    ("ff_seg", ["5", "5"], None),
    ("ff_seg", ["5", "6"], None),
    ("ff_seg", ["5", "10"], None),
    ("ff_seg", ["5", "34"], None),
    ("ff_seg", ["5", "37"], None),
    ("ff_seg", ["5", "45"], None),
    ("ff_seg", ["5", "65"], None),
    ("ff_seg", ["5", "66"], None),
    ("ff_seg", ["17", "46"], None),
    ("ff_seg", ["101", "130"], None),
    ("ff_seg", ["1001", "1040"], None),
    ("ff_seg", ["4001", "4061"], None),
    ("ff_seg", ["99999", "100030"], None),
    ("ff_seg", ["65001", "65037"], None),
    ("ff_seg", ["6", "40"], None),
    ("ff_seg", ["100", "50"], None),
    ("ff_seg", ["300"], None),
    ("ff_seg", [], None),
    ("ff_seg", ["2", "1000"], None),
    ("ff_seg", ["1048000", "1048575"], None),
    ("ff_seg_cap3000", ["5", "20000"], 3000),
    ("ff_seg_cap1000", ["5", "20000"], 1000),
]


def run(binary, args):
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run([os.path.join(REF, binary)] + args, cwd=d, capture_output=True)
    return p.returncode, NORM.sub(r"\1 time: N", p.stdout.decode())


def main():
    out = []
    for binary, args, primemem in CASES:
        rc, text = run(binary, args)
        out.append({"args": args, "primemem": primemem, "rc": rc, "stdout": text})
        print(binary, args, "rc", rc, "lines", text.count("\n"))
    with open(os.path.join(HERE, "seg_cases.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
