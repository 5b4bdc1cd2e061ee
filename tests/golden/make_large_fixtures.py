"""Writes seg_large.json: ff_seg's own stdout for ranges beyond the goldens.

The goldens stop at 2^21.  This script runs the reference program built from
reference/segmentedSieve.C by oracle/Makefile, in a fresh directory per case,
and stores the timing-normalised stdout and exit status:
  * oracle/_ref/ff_seg_cap16g (PRIMEMEM 16 GiB, fits this host's memory) on
    ranges across the search's split at pmapSumLimit = 2097142
    (segmentedSieve.C:905-912), one with a 21-sum first call (thread-split
    quirks);
  * oracle/_ref/ff_seg (uncapped: a 35 GB map) on sums near 3,000,000, where
    the probe values of the wider splits lie above the map and the
    Miller-Rabin branch decides them.
Ranges that start above pmapSumLimit under a capped map are not used: there
the reference's first GenerateThreads call gets sumStart > sumLimit, its
partition arithmetic wraps, and it searches unrelated sums (hundreds of
millions) without finishing.

Run from the repo root after `make -C oracle`:  python tests/golden/make_large_fixtures.py
"""
import json
import os
import re
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref")
NORM = re.compile(r"(Prime|Freudenthal) time: [0-9]+")

CASES = [
    # (binary, args, PRIMEMEM of that binary)
    ("ff_seg_cap16g", ["2097101", "2097200"], 16 << 30),
    ("ff_seg_cap16g", ["2096001", "2099000"], 16 << 30),
    ("ff_seg", ["3000001", "3002000"], None),
]


def main():
    out = []
    for binary, args, primemem in CASES:
        t0 = time.time()
        with tempfile.TemporaryDirectory() as d:
            p = subprocess.run([os.path.join(REF, binary)] + args, cwd=d, capture_output=True)
        text = NORM.sub(r"\1 time: N", p.stdout.decode())
        out.append({"args": args, "primemem": primemem, "rc": p.returncode, "stdout": text})
        print(binary, args, "rc", p.returncode, "lines", text.count("\n"), f"{time.time() - t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "seg_large.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    sys.exit(main())
