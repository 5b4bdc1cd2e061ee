"""ff_sieve end to end on the GPU: stdout byte-identical to ff_seg (golden
hashes of the 65536, 1048576 and 2097152 legs; ff_seg's own stdout for the
edge cases), --dump-map bytes, --map-cap, and the memory run-gate."""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import ffxpy
from conftest import GOLDEN, golden_sha, normalize, run_ffref

pytestmark = pytest.mark.gpu


def ff_sieve(*args, timeout=300):
    print("[case] ff_sieve", *args, flush=True)
    sys.stdout.flush()
    return subprocess.run([ffxpy.FF_SIEVE] + [str(a) for a in args], capture_output=True, timeout=timeout)


@pytest.mark.parametrize("limit,leg", [(65535, 65536), (65536, 65536), (1048575, 1048576), (2097151, 2097152)])
def test_legs_match_golden_hashes(limit, leg):
    p = ff_sieve(5, limit)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    out = normalize(p.stdout.decode())
    assert hashlib.sha256(out.encode()).hexdigest() == golden_sha(leg)


def test_edge_cases_match_ff_seg():
    for c in json.load(open(os.path.join(GOLDEN, "seg_cases.json"))):
        args = (["--primemem", c["primemem"]] if c["primemem"] else []) + c["args"]
        p = ff_sieve(*args)
        assert (p.returncode & 255) == (c["rc"] & 255), (c["args"], p.stderr.decode()[-500:])
        assert normalize(p.stdout.decode()) == c["stdout"], c["args"]


def test_large_sums_match_ff_seg_fixtures():
    # ff_seg's stdout beyond the goldens: across pmapSumLimit under a 16 GiB
    # PRIMEMEM cap (three header lines, the search split in two calls) and
    # near 3,000,000 uncapped (a 35 GB map)
    for c in json.load(open(os.path.join(GOLDEN, "seg_large.json"))):
        args = (["--primemem", c["primemem"]] if c["primemem"] else []) + c["args"]
        p = ff_sieve(*args, timeout=600)
        assert (p.returncode & 255) == (c["rc"] & 255), (c["args"], p.stderr.decode()[-500:])
        assert normalize(p.stdout.decode()) == c["stdout"], c["args"]


def test_map_cap_keeps_stdout():
    # a device map 256x smaller than ff_seg's: Miller-Rabin answers above it
    p = ff_sieve("--map-cap", 16777215, 5, 262143)
    assert p.returncode == 0
    assert hashlib.sha256(normalize(p.stdout.decode()).encode()).hexdigest() == golden_sha(262144)


def test_dump_map_matches_oracle():
    with tempfile.TemporaryDirectory() as d:
        a, b = os.path.join(d, "gpu.bin"), os.path.join(d, "ref.bin")
        p = ff_sieve("--dump-map", a, "--slab-size", "4MiB", 5, 65535)
        assert p.returncode == 0, p.stderr.decode()
        assert run_ffref(["--dump-map", b, 5, 65535]).returncode == 0
        assert open(a, "rb").read() == open(b, "rb").read()


def test_run_gate_refuses_small_budget():
    # backing of a 2 GiB budget cannot hold the 4 GiB map (source/main.cpp:319-334)
    p = ff_sieve("--vram-budget", "2GiB", "--slab-size", "64MiB", 5, 1048575)
    assert p.returncode == 1
    assert b"GATE FAIL" in p.stderr
    assert p.stdout == b""
    # a slab larger than the backing is a validation error (:260-270)
    p = ff_sieve("--vram-budget", "1GiB", 5, 1048575)
    assert p.returncode == 1 and b"--slab-size" in p.stderr
    # the same budget passes when the device map is capped (Miller-Rabin above it)
    p = ff_sieve("--vram-budget", "2GiB", "--slab-size", "64MiB", "--map-cap", 2 ** 30, 5, 65535)
    assert p.returncode == 0
    assert hashlib.sha256(normalize(p.stdout.decode()).encode()).hexdigest() == golden_sha(65536)
