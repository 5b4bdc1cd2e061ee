"""Writes tests/golden/solve_beyond_2p21.txt.gz: the oracle's F(s) for a
seeded sample of odd sums above the goldens' reach (2^21), where the batched
workload (sums below 2^22 + 2^16) and the served stream (sums up to 2^23)
run.

Two strata, seed 20261205 (Python's random.Random):
  2000 odd sums uniform in [2^21, 2^22 + 2^16)
  4000 odd sums uniform in [2^22, 2^23]
Each line is "sum low high" from `oracle/ffref --solve` (low = high = 0 when
the sum is not emitted).  The oracle is pinned to the reference up to 2^21
by the goldens and beyond it by ff_seg's own stdout (seg_large.json); this
sample lets the GPU tests check the plain slot array the batched and served
checks rest on against something other than the GPU.

Run from the repo root (needs oracle/ffref; about a minute on 8 cores):
  python tests/golden/make_solve_fixtures.py
"""
import gzip
import os
import random
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FFREF = os.path.join(ROOT, "oracle", "ffref")
OUT = os.path.join(HERE, "solve_beyond_2p21.txt.gz")
SEED = 20261205


def sample():
    rng = random.Random(SEED)
    sums = set()
    while len(sums) < 2000:
        sums.add(rng.randrange(1 << 21, (1 << 22) + (1 << 16)) | 1)
    top = set()
    while len(top) < 4000:
        s = rng.randrange(1 << 22, (1 << 23) + 1) | 1
        if s <= (1 << 23) and s not in sums:
            top.add(s)
    return sorted(sums | top)


def main():
    sums = sample()
    p = subprocess.run([FFREF, "--solve"], input="\n".join(map(str, sums)).encode(), capture_output=True, check=True)
    rows = [l for l in p.stdout.decode().split("\n") if l]
    assert len(rows) == len(sums)
    with gzip.open(OUT, "wt") as f:
        f.write("\n".join(rows) + "\n")
    print(f"wrote {OUT}: {len(rows)} sums, {sum(1 for r in rows if r.split()[1] != '0')} emitted")


if __name__ == "__main__":
    main()
