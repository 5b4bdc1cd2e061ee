"""Sieve kernel parity (GPU): byte-exact against the oracle's map, per-slab
ranges after test/source/slab_cmp.cpp:232-265, popcount == pi(M) - 1."""
import math
import os
import sys
import tempfile

import numpy as np
import pytest

import ffxpy
from conftest import run_ffref

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ffxpy.init(0)


def oracle_map(sum_limit):
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "m.bin")
        assert run_ffref(["--dump-map", f, 5, sum_limit]).returncode == 0
        return np.fromfile(f, dtype=np.uint8)


def small_primes_tail(lo, hi):
    return sum(1 for n in range(lo, hi + 1) if n > 1 and all(n % p for p in range(2, math.isqrt(n) + 1)))


@pytest.mark.parametrize("sum_limit", [5, 7, 100, 8191, 65535, 131071])
def test_full_map_matches_oracle(sum_limit):
    ref = oracle_map(sum_limit)
    M = len(ref) * 16 - 1
    say("full map", sum_limit, M)
    buf = ffxpy.sieve_map(M)
    got = buf.to_host(len(ref))
    assert np.array_equal(got, ref)


SLAB_RANGES = [(0, 65536), (0, 524288), (0, 1048576), (500000, 524288), (100000, 300000), (0, 256),
               (524288, 1048576), (524272, 524400), (1200, 2000), (0, 2000000),
               # slabs straddling the kernel's 64 KiB segment edges, unaligned ends
               (1048576 - 16, 1048576 + 33), (123456 * 16, 123456 * 16 + 999999), (16, 17)]


def test_slabs_match_map():
    ref = oracle_map(65535)
    M = len(ref) * 16 - 1
    plan = ffxpy.SievePlan(M)
    lib = ffxpy.load()
    for lo, hi in SLAB_RANGES:
        say("slab", lo, hi)
        nb = (hi + 15) // 16 - lo // 16
        slab = ffxpy.DeviceBuffer(nb)
        ffxpy.check(lib.ffx_sieve_slab_async(plan.h, lo, hi, slab.ptr, None))
        ffxpy.check(lib.ffx_device_sync())
        assert np.array_equal(slab.to_host(nb), ref[lo // 16: lo // 16 + nb]), (lo, hi)
    assert lib.ffx_sieve_slab_async(plan.h, 8, 100, None, None) == -22


def test_map_grown_in_ranges_equals_full_map():
    ref = oracle_map(131071)
    M = len(ref) * 16 - 1
    plan = ffxpy.SievePlan(M)
    lib = ffxpy.load()
    buf = ffxpy.DeviceBuffer((len(ref) + 255) & ~255)
    cuts = [0, 1, 65536 * 3 + 7, 10 ** 6 + 3, 40 * 10 ** 6 + 1, len(ref)]
    for a, b in zip(cuts, cuts[1:]):
        ffxpy.check(lib.ffx_sieve_map_range_async(plan.h, buf.ptr, a, b, None))
    ffxpy.check(lib.ffx_device_sync())
    assert np.array_equal(buf.to_host(len(ref)), ref)


# (segment bytes, hold primes, lane carry): every kernel variant, with its
# carried lane primes covering none, part or all of the 3,500 sieving primes
# of a 1 GiB map (1024 threads per workgroup), so both the carried strikes
# and the per-segment first multiples run
CARRY_CASES = [(32768, 0, 0), (32768, 0, 1), (32768, 0, 64), (114688, 1, 0), (114688, 1, 2), (114688, 1, 64),
               (114688, 2, 0), (114688, 2, 3), (114688, 2, 64)]


@pytest.mark.parametrize("seg,hold,carry", CARRY_CASES)
def test_lane_carry_variants_match_oracle(seg, hold, carry):
    ref = oracle_map(131071)
    M = len(ref) * 16 - 1
    say("carry variant", seg, hold, carry, M)
    plan = ffxpy.SievePlan(M)
    lib = ffxpy.load()
    for param, v in ((ffxpy.SIEVE_SEGMENT_BYTES, seg), (ffxpy.SIEVE_HOLD_PRIMES, hold), (ffxpy.SIEVE_LANE_CARRY, carry)):
        ffxpy.check(lib.ffx_sieve_plan_set(plan.h, param, v))
    buf = ffxpy.sieve_map(M, plan)
    assert np.array_equal(buf.to_host(len(ref)), ref)
    # ranges grown on the same plan (runs that start mid-map)
    buf2 = ffxpy.DeviceBuffer((len(ref) + 255) & ~255)
    cuts = [0, 3 * 65536 + 5, 10 ** 7 + 1, len(ref)]
    for a, b in zip(cuts, cuts[1:]):
        ffxpy.check(lib.ffx_sieve_map_range_async(plan.h, buf2.ptr, a, b, None))
    ffxpy.check(lib.ffx_device_sync())
    assert np.array_equal(buf2.to_host(len(ref)), ref)


def test_plan_from_caller_primes():
    ref = oracle_map(65535)
    M = len(ref) * 16 - 1
    lim = math.isqrt(M)
    s = np.ones(lim + 1, dtype=bool)
    s[:2] = False
    for p in range(2, math.isqrt(lim) + 1):
        if s[p]:
            s[p * p::p] = False
    primes = np.nonzero(s)[0].astype(np.uint32)  # includes 2, as the reference's test list does
    plan = ffxpy.SievePlan(M, primes=primes)
    buf = ffxpy.sieve_map(M, plan)
    assert np.array_equal(buf.to_host(len(ref)), ref)
    with pytest.raises(ffxpy.FfxError):
        ffxpy.SievePlan(M, primes=primes[: len(primes) // 2])


@pytest.mark.parametrize("k,pi", [(32, 203280221), (36, 2874398515), (38, 10866266172)])
def test_popcount_is_pi(k, pi):
    # maps of the 262143 / 1048575 / 2097151 legs: M = 2^k + 15
    M = 2 ** k + 15
    say("popcount", k)
    buf = ffxpy.sieve_map(M)
    nb = ffxpy.map_bytes_for(M)
    total = 0
    step = 1 << 30
    for off in range(0, nb, step):
        n = min(step, nb - off)
        total += int(np.bitwise_count(buf.to_host(n, offset=off)).sum())
    assert total == pi + small_primes_tail(2 ** k + 1, M) - 1


def map_hashes():
    """sumLimit -> (map bytes, sha256) of the reference program's own
    --dump-map (tests/golden/make_map_hashes.py)."""
    out = {}
    with open(os.path.join(os.path.dirname(__file__), "golden", "map_sha256.txt")) as f:
        for line in f:
            L, n, h = line.split()
            out[int(L)] = (int(n), h)
    return out


@pytest.mark.parametrize("sum_limit", [1048575, 2097151])
def test_map_sha256_matches_reference_dump(sum_limit):
    # the whole 4 GiB / 16 GiB device map of odd sums [5, 2^20) and
    # [5, 2^21) hashes to the reference's own dump of it
    # (scripts/verify.sh:156-173 makes the same check)
    import hashlib
    n, want = map_hashes()[sum_limit]
    M = {1048575: 68719476751, 2097151: 274877906959}[sum_limit]
    assert ffxpy.map_bytes_for(M) == n
    say("map sha256", sum_limit)
    buf = ffxpy.sieve_map(M)
    h = hashlib.sha256()
    step = 1 << 29
    for off in range(0, n, step):
        h.update(buf.to_host(min(step, n - off), offset=off))
    buf.close()
    assert h.hexdigest() == want
