"""ff_sieve end to end on the GPU: stdout byte-identical to ff_seg (golden
hashes of the 65536, 1048576 and 2097152 legs; ff_seg's own stdout for the
edge cases), --dump-map bytes, --map-cap, the memory run-gate and the host
tier."""
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


def test_host_tier_dump_map_matches_oracle():
    # a 15 MiB backing (80 MiB budget - 1 MiB scratch - 64 MiB headroom) and a
    # 64 MiB map: the host tier holds it, streamed through three 4 MiB device
    # slab buffers (source/pull_scheduler.cpp:164-200)
    with tempfile.TemporaryDirectory() as d:
        a, b = os.path.join(d, "gpu.bin"), os.path.join(d, "ref.bin")
        p = ff_sieve("--dump-map", a, "--vram-budget", "80MiB", "--scratch", "1MiB", "--slab-size", "4MiB",
                     "--host-tier-cap", "1GiB", 5, 131071)
        assert p.returncode == 0, p.stderr.decode()[-2000:]
        assert b"host tier" in p.stderr and b"GATE PASS" in p.stderr
        assert run_ffref(["--dump-map", b, 5, 131071]).returncode == 0
        assert os.path.getsize(a) == os.path.getsize(b) > 4 * (15 << 20)
        assert open(a, "rb").read() == open(b, "rb").read()
        # without the host tier the same budget fails the gate
        p = ff_sieve("--dump-map", a, "--vram-budget", "80MiB", "--scratch", "1MiB", "--slab-size", "4MiB", 5, 131071)
        assert p.returncode == 1 and b"GATE FAIL" in p.stderr


def test_host_tier_dump_of_the_2p20_map_matches_reference_hash():
    # the 4 GiB map of odd sums [5, 2^20) through a 2 GiB budget: streamed
    # into host memory and hashed as it is written (a pipe, not a file), equal
    # to the hash of the reference program's own --dump-map
    import threading
    n, want = map_hashes()[1048575]
    with tempfile.TemporaryDirectory() as d:
        fifo = os.path.join(d, "map.pipe")
        os.mkfifo(fifo)
        h = hashlib.sha256()
        got = [0]

        def drain():
            with open(fifo, "rb") as f:
                while True:
                    chunk = f.read(1 << 24)
                    if not chunk:
                        break
                    h.update(chunk)
                    got[0] += len(chunk)

        t = threading.Thread(target=drain)
        t.start()
        try:
            p = ff_sieve("--dump-map", fifo, "--vram-budget", "2GiB", "--slab-size", "256MiB", "--host-tier-cap",
                         "4GiB", 5, 1048575, timeout=300)
        finally:
            if t.is_alive():  # the writer never opened the pipe: open it once so the reader ends
                with open(fifo, "wb"):
                    pass
            t.join(60)
        assert p.returncode == 0, p.stderr.decode()[-2000:]
        assert b"host tier" in p.stderr
        assert got[0] == n and h.hexdigest() == want


def test_host_tier_search_keeps_stdout():
    # the 2 GiB budget that fails the 2^20 leg alone passes with a host tier:
    # the search keeps a 1.6 GiB device map and Miller-Rabin answers above it
    p = ff_sieve("--vram-budget", "2GiB", "--slab-size", "64MiB", "--host-tier-cap", "4GiB", 5, 1048575)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    assert b"GATE PASS" in p.stderr and b"host tier" in p.stderr
    assert hashlib.sha256(normalize(p.stdout.decode()).encode()).hexdigest() == golden_sha(1048576)
    # a 15 MiB backing under the 2^18 leg (a 256 MiB map)
    p = ff_sieve("--vram-budget", "80MiB", "--scratch", "1MiB", "--slab-size", "4MiB", "--host-tier-cap", "1GiB",
                 5, 262143)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    assert hashlib.sha256(normalize(p.stdout.decode()).encode()).hexdigest() == golden_sha(262144)


def map_hashes():
    out = {}
    with open(os.path.join(GOLDEN, "map_sha256.txt")) as f:
        for line in f:
            L, n, h = line.split()
            out[int(L)] = (int(n), h)
    return out
