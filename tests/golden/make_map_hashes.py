"""Writes tests/golden/map_sha256.txt: the SHA-256 of the reference's own
bitmap for the C2 and C3 workloads (odd sums [5, 1048575] and
[5, 2097151]: 4 GiB and 16 GiB maps), the check the reference's
scripts/verify.sh:156-173 makes of a GPU map.

The maps come from the reference program itself, built from its source by
oracle/Makefile (oracle/_ref/ff_seg), through its --dump-map option
(reference/segmentedSieve.C:828-843, :883-902).  Each dump is written into a
named pipe and hashed as it streams, so no file of the map's size is kept.
Lines: "sumLimit mapBytes sha256".

Run from the repo root, in the container that holds the reference (a few
minutes; the 16 GiB map needs that much free memory):
  python tests/golden/make_map_hashes.py
"""
import hashlib
import os
import subprocess
import tempfile
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FF_SEG = os.path.join(ROOT, "oracle", "_ref", "ff_seg")
OUT = os.path.join(HERE, "map_sha256.txt")
LIMITS = [1048575, 2097151]


def dump_hash(sum_limit):
    with tempfile.TemporaryDirectory() as d:
        fifo = os.path.join(d, "map.fifo")
        os.mkfifo(fifo)
        res = {}

        def reader():
            h = hashlib.sha256()
            n = 0
            with open(fifo, "rb") as f:
                while True:
                    b = f.read(1 << 24)
                    if not b:
                        break
                    h.update(b)
                    n += len(b)
            res["sha"], res["bytes"] = h.hexdigest(), n

        t = threading.Thread(target=reader)
        t.start()
        work = os.path.join(d, "work")
        os.mkdir(work)
        p = subprocess.run([FF_SEG, "--dump-map", fifo, "5", str(sum_limit)], cwd=work, capture_output=True)
        t.join()
        assert p.returncode == 0, p.stdout.decode()
    return res["bytes"], res["sha"]


def main():
    lines = []
    for L in LIMITS:
        n, sha = dump_hash(L)
        print(f"{L} {n} {sha}", flush=True)
        lines.append(f"{L} {n} {sha}\n")
    with open(OUT, "w") as f:
        f.writelines(lines)


if __name__ == "__main__":
    main()
