import gzip
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ff-gpu_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
FFREF = os.path.join(ROOT, "oracle", "ffref")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

NORM = re.compile(r"(Prime|Freudenthal) time: [0-9]+")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def normalize(text: str) -> str:
    """Timing digits -> N, as the reference's goldens are stored."""
    return NORM.sub(r"\1 time: N", text)


def golden_sha(leg: int) -> str:
    with open(os.path.join(GOLDEN, "SHA256SUMS")) as f:
        for line in f:
            h, name = line.split()
            if name == f"out_ff_seg_{leg}.txt":
                return h
    raise KeyError(leg)


def golden_solutions():
    """(sum, low, tags) of every solution with sum <= 1048576."""
    rows = []
    with gzip.open(os.path.join(GOLDEN, "solutions_1048576.txt.gz"), "rt") as f:
        for line in f:
            s, lo, t = line.split()
            rows.append((int(s), int(lo), int(t)))
    return rows


def golden_headers():
    out = {}
    with open(os.path.join(GOLDEN, "headers.txt")) as f:
        for line in f:
            leg, h1, h2 = line.rstrip("\n").split("\t")
            out[int(leg)] = (h1, h2)
    return out


def tags_to_flags(t: int) -> int:
    """Fixture tag code -> FFX_TAG_* bits."""
    return ((t & 1) << 28) | ((t & 2) << 28) | ((t & 4) << 28) | ((t & 8) << 28)


def expected_text(leg: int) -> str:
    """ff_seg stdout (timing-normalised) for `5 leg`, rebuilt from fixtures."""
    from ffxpy import format_line
    h1, h2 = golden_headers()[leg]
    lines = [h1 + "\n", h2 + "\n"]
    n = 0
    for s, lo, t in golden_solutions():
        if s > leg:
            break
        n += 1
        lines.append(format_line(n, s, lo, s - lo, tags_to_flags(t)))
    lines.append("Prime time: N μs\nFreudenthal time: N μs\n")
    return "".join(lines)


@pytest.fixture(scope="session")
def ffref():
    if not os.path.exists(FFREF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ffref"], check=True, capture_output=True)
    return FFREF


def run_ffref(args, stdin=None, timeout=600):
    p = subprocess.run([FFREF] + [str(a) for a in args], input=stdin, capture_output=True, timeout=timeout)
    return p
