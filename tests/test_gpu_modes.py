"""Batched (--batch) and served (--serve) ff_sieve on the GPU.

Every range of a batch and every request of a served stream must print
exactly what a plain `ff_sieve <sumStart> <sumLimit>` (and ff_seg) prints for
it alone: checked against the oracle (oracle/ffref, ff_seg-identical stdout)
on small sums, and at full size (the 4096-range batch, the 20000-request
served stream) range by range against one plain slot array, itself checked
against the oracle's answers for a seeded sample of sums above 2^21."""
import gzip
import hashlib
import json
import os
import random
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


def summary(p):
    assert p.returncode == 0, p.stderr.decode()[-3000:]
    return json.loads(p.stdout.decode().strip().splitlines()[-1])


def oracle_text(a, b):
    p = run_ffref([a, b], timeout=120)
    assert p.returncode == 0
    return normalize(p.stdout.decode())


EDGE_RANGES = [(5, 5), (4, 6), (1, 100), (100, 50), (7, 2627), (2, 9), (1000, 1003), (65535, 65535),
               (65536, 65600), (60001, 70000), (60001, 70000), (61001, 65003), (17, 17), (131000, 131071)]


def seeded_ranges(seed, n, top):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        a = rng.randrange(5, top)
        w = int(2 ** rng.uniform(2, 13))
        out.append((a, min(top, a + w - 1)))
    return out


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_ranges_match_oracle(graph):
    ranges = EDGE_RANGES + seeded_ranges(11, 34, 1 << 17)
    with tempfile.TemporaryDirectory() as d:
        rf = os.path.join(d, "ranges.txt")
        with open(rf, "w") as f:
            f.writelines(f"{a} {b}\n" for a, b in ranges)
        s = summary(ff_sieve("--batch", "--opt", f"ranges={rf}", "--opt", f"out_dir={d}", "--opt", f"graph={graph}",
                             "--opt", "reps=2"))
        assert s["ranges"] == len(ranges) and s["graph"] == bool(graph)
        lines = 0
        for i, (a, b) in enumerate(ranges):
            got = normalize(open(os.path.join(d, f"range_{i:05d}.txt")).read())
            assert got == oracle_text(a, b), (i, a, b)
            lines += got.count(") sum =")
        assert s["solution_lines"] == lines


def fixture_beyond_2p21():
    """The oracle's F(s) for 6000 seeded sums in [2^21, 2^23]
    (tests/golden/make_solve_fixtures.py): {sum: low}, 0 = not emitted."""
    rows = [l.split() for l in gzip.open(os.path.join(GOLDEN, "solve_beyond_2p21.txt.gz"), "rt") if l.strip()]
    return {int(s): int(lo) for s, lo, _ in rows}


def check_plain_slots_against_oracle(path, max_sum):
    """The plain slot array a verified run wrote (u32 per odd sum from 5 to
    the largest sum any range asks for, at most max_sum) against the
    oracle's sample: the check of every range rests on it."""
    slots = np.fromfile(path, dtype="<u4")
    top = 5 + 2 * (len(slots) - 1)
    assert max_sum - 2048 < top <= max_sum, (top, max_sum)
    max_sum = top
    fx = {s: lo for s, lo in fixture_beyond_2p21().items() if s <= max_sum}
    assert len(fx) > 1000
    got = {s: int(slots[(s - 5) // 2]) & ffxpy.SLOT_LOW_MASK for s in fx}
    bad = [s for s in fx if got[s] != fx[s]]
    assert not bad, bad[:10]
    return len(fx), sum(1 for v in fx.values() if v)


def test_batch_c4_every_range_matches_plain_slots():
    # the full batched workload (4096 seeded ranges, one graph-captured
    # launch): every range's stdout digest against the same range formatted
    # from one plain slot array over [5, max sum], which in turn matches the
    # oracle's answers for the seeded sums above 2^21
    with tempfile.TemporaryDirectory() as d:
        slots = os.path.join(d, "plain_slots.u32")
        s = summary(ff_sieve("--batch", "--opt", "verify=1", "--opt", f"slots_out={slots}", "--opt",
                             f"hashes={os.path.join(d, 'digests.txt')}", timeout=600))
        assert s["ranges"] == 4096 and s["graph"] is True
        assert s["parity_checked"] == 4096 and s["parity_mismatches"] == 0, s
        rows = open(os.path.join(d, "digests.txt")).read().splitlines()
        assert len(rows) == 4096 and sum(int(r.split()[3]) for r in rows) == s["solution_lines"]
        max_sum = max(int(r.split()[2]) for r in rows)
        n, emitted = check_plain_slots_against_oracle(slots, max_sum if max_sum % 2 else max_sum - 1)
        print(f"[case] C4: 4096 ranges match the plain slots; {n} oracle sums ({emitted} emitted) match", flush=True)


def test_serve_c5_every_request_matches_plain_slots():
    # the served stream as configured (seed 20261204: 20000 Poisson requests
    # at 2000/s, sumLimit up to 2^23, widths 2^6..2^18, map grown on demand):
    # every request's stdout digest, taken as it completes in the mix,
    # against the request formatted from one plain slot array over [5, 2^23]
    # (itself checked against the oracle's sample)
    with tempfile.TemporaryDirectory() as d:
        slots = os.path.join(d, "plain_slots.u32")
        s = summary(ff_sieve("--serve", "--opt", "verify=1", "--opt", f"slots_out={slots}", timeout=900))
        assert s["requests"] == 20000 and s["map_growths"] > 1
        assert s["parity_checked"] == 20000 and s["parity_mismatches"] == 0, s
        n, emitted = check_plain_slots_against_oracle(slots, (1 << 23) - 1)
        print(f"[case] C5: 20000 requests match the plain slots; {n} oracle sums ({emitted} emitted) match",
              flush=True)


def test_serve_requests_match_oracle_alone():
    # each request's output in the mix (continuous batching over two
    # streams, shared result slots) equals its output alone
    count = 60
    with tempfile.TemporaryDirectory() as d:
        s = summary(ff_sieve("--serve", "--opt", f"count={count}", "--opt", "seed=7", "--opt", "lim_lo=1024",
                             "--opt", "lim_hi=131071", "--opt", "wlo=2", "--opt", "whi=13", "--opt", "rate=3000",
                             "--opt", "batch_sums=512", "--opt", "streams=2", "--opt", f"out_dir={d}"))
        assert s["requests"] == count and s["batches"] >= 2
        reqs = [tuple(map(int, l.split()[:2])) for l in subprocess.run(
            [os.path.join(os.path.dirname(ffxpy.FF_SIEVE), "ff_hostcheck"), "served", "7", str(count), "3000",
             "1024", "131071", "2", "13"], capture_output=True, text=True, check=True).stdout.splitlines()]
        assert len(reqs) == count
        for i, (a, b) in enumerate(reqs):
            got = normalize(open(os.path.join(d, f"request_{i:05d}.txt")).read())
            assert got == oracle_text(a, b), (i, a, b)


def test_serve_mix_matches_plain_on_large_sums():
    # requests up to sumLimit 2^23 (capped PRIMEMEM headers included), a
    # device map bounded by --map-cap (8 GiB: grown while batches run); each
    # against a plain run
    count = 24
    with tempfile.TemporaryDirectory() as d:
        s = summary(ff_sieve("--map-cap", 2 ** 37, "--serve", "--opt", "verify=1", "--opt", f"count={count}", "--opt", "seed=9",
                             "--opt", "wlo=6", "--opt", "whi=12", "--opt", "rate=500", "--opt", f"out_dir={d}",
                             timeout=600))
        assert s["requests"] == count
        reqs = [tuple(map(int, l.split()[:2])) for l in subprocess.run(
            [os.path.join(os.path.dirname(ffxpy.FF_SIEVE), "ff_hostcheck"), "served", "9", str(count), "500",
             str(1 << 18), str(1 << 23), "6", "12"], capture_output=True, text=True, check=True).stdout.splitlines()]
        for i in random.Random(3).sample(range(count), 6):
            a, b = reqs[i]
            plain = ff_sieve("--map-cap", 2 ** 37, a, b, timeout=600)
            assert plain.returncode == 0
            got = normalize(open(os.path.join(d, f"request_{i:05d}.txt")).read())
            assert got == normalize(plain.stdout.decode()), (i, a, b)


def test_autotune_writes_a_tune_file_that_keeps_stdout():
    # every candidate is checked against the default configuration inside
    # the autotuner; the file it writes drives a plain run to the golden
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "tune.json")
        p = ff_sieve("--autotune", "--opt", "sum_limit=65535", "--opt", "tune_reps=1", "--opt", f"out={f}")
        assert p.returncode == 0, p.stderr.decode()[-2000:]
        t = json.load(open(f))
        assert t["sieve_segment_bytes"] in (32768, 114688) and 1 <= t["search_blocks_per_cu"] <= 16
        assert len(t["trials"]["search"]) == 36 and len(t["trials"]["sieve"]) == 8
        p = ff_sieve("--tune", f, 5, 65535)
        assert p.returncode == 0
        assert hashlib.sha256(normalize(p.stdout.decode()).encode()).hexdigest() == golden_sha(65536)
