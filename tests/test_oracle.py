"""The CPU oracle (oracle/ffref) against the reference's own outputs and
against mathematics.  No GPU."""
import hashlib
import json
import math
import os
import random
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, REF_DIR, expected_text, golden_sha, normalize, run_ffref


def sha(text: str) -> str:
    return hashlib.sha256(text.encode()).hexdigest()


def test_fixture_text_matches_reference_hashes():
    # the fixtures rebuild the goldens exactly (pins the fixtures themselves)
    for leg in (65536, 131072, 262144, 524288, 1048576):
        assert sha(expected_text(leg)) == golden_sha(leg), leg


@pytest.mark.parametrize("leg", [65536, 131072, 262144])
def test_oracle_stdout_matches_golden(ffref, leg):
    p = run_ffref([5, leg])
    assert p.returncode == 0
    out = normalize(p.stdout.decode())
    assert sha(out) == golden_sha(leg)


@pytest.mark.skipif(os.environ.get("FF_SLOW") != "1", reason="set FF_SLOW=1 (minutes of CPU)")
@pytest.mark.parametrize("leg", [524288, 1048576, 2097152])
def test_oracle_stdout_matches_golden_large(ffref, leg):
    p = run_ffref([5, leg], timeout=3600)
    assert p.returncode == 0
    assert sha(normalize(p.stdout.decode())) == golden_sha(leg)


def test_oracle_edge_cases_match_ff_seg(ffref):
    # tiny ranges, positional quirks and capped geometry, captured from the
    # reference program itself (tests/golden/make_seg_fixtures.py)
    cases = json.load(open(os.path.join(GOLDEN, "seg_cases.json")))
    for c in cases:
        args = (["--primemem", c["primemem"]] if c["primemem"] else []) + c["args"]
        p = run_ffref(args)
        assert (p.returncode & 255) == (c["rc"] & 255), c["args"]
        assert normalize(p.stdout.decode()) == c["stdout"], c["args"]


def _primes_upto(n):
    s = np.ones(n + 1, dtype=bool)
    s[:2] = False
    for p in range(2, int(n ** 0.5) + 1):
        if s[p]:
            s[p * p::p] = False
    return s


def _map_bits(m: np.ndarray) -> np.ndarray:
    """bit of odd value 2j+1 at index j (MSB of each byte first)."""
    return np.unpackbits(m)


def test_oracle_bitmap_brute_force(ffref):
    # bit(n) == isprime(n) for every odd n <= M on a small map
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "m.bin")
        assert run_ffref(["--dump-map", f, 5, 8191]).returncode == 0  # M = 4194319
        m = np.fromfile(f, dtype=np.uint8)
    M = len(m) * 16 - 1
    bits = _map_bits(m)
    isp = _primes_upto(M)
    odd = np.arange(1, M + 1, 2)
    assert np.array_equal(bits[: len(odd)].astype(bool), isp[odd])


def test_oracle_bitmap_pi_2_32(ffref):
    # popcount of the map for sumLimit 262143 (M = 2^32 + 15) = pi(M) - 1
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "m.bin")
        assert run_ffref(["--dump-map", f, 5, 262143]).returncode == 0
        m = np.fromfile(f, dtype=np.uint8)
    M = len(m) * 16 - 1
    assert M == 2 ** 32 + 15
    pi_2_32 = 203280221  # OEIS A007053
    tail = sum(1 for n in range(2 ** 32 + 1, M + 1) if all(n % p for p in range(2, int(n ** 0.5) + 1)))
    assert int(np.bitwise_count(m).sum()) == pi_2_32 + tail - 1


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "ff_seg")), reason="reference binary not built")
def test_oracle_map_matches_ff_seg_dump(ffref):
    with tempfile.TemporaryDirectory() as d:
        a, b = os.path.join(d, "a.bin"), os.path.join(d, "b.bin")
        assert run_ffref(["--dump-map", a, 5, 65535]).returncode == 0
        p = subprocess.run([os.path.join(REF_DIR, "ff_seg"), "--dump-map", b, "5", "65535"], cwd=d,
                           capture_output=True)
        assert p.returncode == 0
        assert open(a, "rb").read() == open(b, "rb").read()


def _trial_prime(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    r = math.isqrt(n)
    p = 3
    while p <= r:
        if n % p == 0:
            return False
        p += 2
    return True


def test_oracle_miller_rabin(ffref):
    rng = random.Random(7)
    ns = [rng.randrange(10 ** 6, 10 ** 11) | 1 for _ in range(300)]
    ns += [p for p in (1000000007, 999999999989, 2 ** 31 - 1, 4294967311)]
    # strong pseudoprimes to the first 4/5/6 bases: the base count by
    # magnitude must still call them composite
    ns += [3215031751, 2152302898747, 3474749660383]
    # products of two primes near the sqrt (the hard cases for MR)
    ns += [1000003 * 1000033, 999983 * 999979, 4294967291 * 3]
    p = run_ffref(["--mr"], stdin="\n".join(map(str, ns)).encode())
    got = {int(a): int(b) for a, b in (l.split() for l in p.stdout.decode().split("\n") if l)}
    for n in ns:
        assert got[n] == int(_trial_prime(n)), n


def test_oracle_solve_matches_goldens(ffref):
    # the pure per-sum mode agrees with the golden solutions on a sample,
    # with a bitmap far smaller than ff_seg's (Miller-Rabin above 2^20)
    from conftest import golden_solutions
    sol = {s: (lo, t) for s, lo, t in golden_solutions()}
    rng = random.Random(3)
    sums = sorted(set([rng.randrange(5, 1048576) | 1 for _ in range(300)] + list(sol)[-40:]))
    p = run_ffref(["--solve", "--map-cap", 1 << 20], stdin="\n".join(map(str, sums)).encode())
    rows = [tuple(map(int, l.split())) for l in p.stdout.decode().split("\n") if l]
    assert len(rows) == len(sums)
    for s, lo, hi in rows:
        if s in sol:
            assert lo == sol[s][0] and hi == s - lo, s
        else:
            assert lo == 0, s


def test_oracle_solve_matches_ff_seg_beyond_the_goldens(ffref):
    # ff_seg's own stdout (tests/golden/seg_large.json, written by
    # make_large_fixtures.py from the reference program) for ranges across
    # pmapSumLimit under a 16 GiB PRIMEMEM cap and near 3,000,000: the
    # oracle's per-sum answers over a 1 GiB map, where Miller-Rabin decides
    # most probes, give exactly the printed lines
    import re
    line = re.compile(r"^\s*\d+\) sum =\s*(\d+), product =\s*\d+,  low term =\s*(\d+)")
    cases = json.load(open(os.path.join(GOLDEN, "seg_large.json")))
    assert len(cases) == 3
    for c in cases:
        a, b = map(int, c["args"])
        want = {int(m.group(1)): int(m.group(2)) for m in (line.match(l) for l in c["stdout"].split("\n")) if m}
        sums = list(range(a | 1, b + 1, 2))
        p = run_ffref(["--solve", "--map-cap", 2 ** 30 - 1], stdin="\n".join(map(str, sums)).encode())
        got = {int(s): int(lo) for s, lo, _ in (l.split() for l in p.stdout.decode().split("\n") if l) if lo != "0"}
        assert got == want, c["args"]
        assert want  # every case prints solutions


def test_solve_fixture_beyond_2p21_is_the_oracles(ffref):
    # tests/golden/solve_beyond_2p21.txt.gz (make_solve_fixtures.py): a
    # seeded slice of it recomputed by the current oracle
    import gzip
    import random
    rows = [l.split() for l in gzip.open(os.path.join(GOLDEN, "solve_beyond_2p21.txt.gz"), "rt") if l.strip()]
    assert len(rows) == 6000
    assert sum(1 for r in rows if r[1] != "0") > 300
    pick = random.Random(9).sample(rows, 96)
    p = run_ffref(["--solve"], stdin="\n".join(r[0] for r in pick).encode())
    got = [l.split() for l in p.stdout.decode().splitlines() if l]
    assert got == pick
