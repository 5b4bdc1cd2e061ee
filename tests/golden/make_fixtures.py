"""Writes the golden fixtures under tests/golden from the reference's goldens.

Reads /root/reference/goldens/out_ff_seg_1048576.txt (ff_seg stdout for
sums [5, 1048576]) and writes solutions_1048576.txt.gz: one "sum low tags"
line per solution line, tags being the sum of 1 (low prime), 2 (low 2^k),
4 (high prime), 8 (high 2^k).  Every smaller leg's solution block is a
prefix of this one (the per-sum output does not depend on sumLimit), so the
file pins the solution lines of legs 65536..1048576; the header lines are
stored per leg in headers.txt.  SHA256SUMS carries the reference's hashes of
the timing-normalised stdout of all six legs, the 2097152 leg included.

Run from the repo root:  python tests/golden/make_fixtures.py
"""
import gzip
import os
import re

REF = "/root/reference/goldens"
HERE = os.path.dirname(os.path.abspath(__file__))
LINE = re.compile(r"^\s*(\d+)\) sum =\s*(\d+), product =\s*(\d+),  low term =\s*(\d+)(.*?)high term =\s*(\d+)(.*)$")


def tags_of(low_tail: str, high_tail: str) -> int:
    t = 0
    if "(prime)" in low_tail:
        t |= 1
    elif "(2^" in low_tail:
        t |= 2
    if "(prime)" in high_tail:
        t |= 4
    elif "(2^" in high_tail:
        t |= 8
    return t


def main():
    rows = []
    with open(os.path.join(REF, "out_ff_seg_1048576.txt")) as f:
        for line in f:
            m = LINE.match(line)
            if m:
                rows.append((int(m.group(2)), int(m.group(4)), tags_of(m.group(5), m.group(7))))
    with gzip.open(os.path.join(HERE, "solutions_1048576.txt.gz"), "wt") as g:
        for s, lo, t in rows:
            g.write(f"{s} {lo} {t}\n")
    with open(os.path.join(HERE, "headers.txt"), "w") as h:
        for leg in (65536, 131072, 262144, 524288, 1048576):
            with open(os.path.join(REF, f"out_ff_seg_{leg}.txt")) as f:
                h.write(f"{leg}\t{f.readline().rstrip()}\t{f.readline().rstrip()}\n")
    print(f"{len(rows)} solutions written")


if __name__ == "__main__":
    main()
