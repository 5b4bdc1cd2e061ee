"""Writes seg_large.json: ff_seg's own stdout for ranges of large sums.

The goldens stop at 2^21.  The batched and served workloads reach 2^22+2^16
and 2^23, where most probe values lie above ff_seg's (capped) map and the
Miller-Rabin branch decides them.  This script runs the reference program
built from reference/segmentedSieve.C by oracle/Makefile with its PRIMEMEM
cap set to 16 GiB (oracle/_ref/ff_seg_cap16g: fits this host's memory; the
cap changes only the header lines and the map size, not which lines are
printed) on short ranges near 2^22 and 2^23 and across the search's split
at pmapSumLimit (segmentedSieve.C:905-912), in a fresh directory per case,
and stores the timing-normalised stdout and exit status.

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
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "ff_seg_cap16g")
PRIMEMEM = 16 << 30
NORM = re.compile(r"(Prime|Freudenthal) time: [0-9]+")

CASES = [
    ["2097101", "2097200"],   # straddles pmapSumLimit = 2097142: a 21-sum first call (thread-split quirks)
    ["2096001", "2099000"],   # straddles pmapSumLimit, 1500 odd sums
    ["4194001", "4194300"],   # top of the batched workload's sums
    ["8388541", "8388607"],   # top of the served workload's sums (2^23 - 1)
]


def main():
    out = []
    for args in CASES:
        t0 = time.time()
        with tempfile.TemporaryDirectory() as d:
            p = subprocess.run([REF] + args, cwd=d, capture_output=True)
        text = NORM.sub(r"\1 time: N", p.stdout.decode())
        out.append({"args": args, "primemem": PRIMEMEM, "rc": p.returncode, "stdout": text})
        print(args, "rc", p.returncode, "lines", text.count("\n"), f"{time.time() - t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "seg_large.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    sys.exit(main())
