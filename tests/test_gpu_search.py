"""Search kernel parity (GPU) through the C ABI: golden solutions, the
oracle's pure per-sum answers (also with a small map, so Miller-Rabin
decides most probes), the reference-compatible record entry point and the
ordered compaction."""
import random
import sys

import numpy as np
import pytest

import ffxpy
from conftest import golden_solutions, run_ffref, tags_to_flags

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ffxpy.init(0)


@pytest.fixture(scope="module")
def c1_map():
    M = 268435471  # ff_seg geometry of 5..65535
    return ffxpy.sieve_map(M), M


def test_c1_solutions_match_goldens(c1_map):
    dmap, M = c1_map
    say("search of odd sums 5..65535")
    got = ffxpy.search(dmap, M, 5, 65535)
    want = [(s, lo, tags_to_flags(t)) for s, lo, t in golden_solutions() if s <= 65535]
    assert len(got) == len(want) == 2357
    assert [(int(r["sum"]), int(r["low"]), int(r["tags"])) for r in got] == want
    assert all(int(r["high"]) == int(r["sum"]) - int(r["low"]) for r in got)


def test_subranges_and_tiny_ranges(c1_map):
    dmap, M = c1_map
    gold = {s: lo for s, lo, _ in golden_solutions() if s <= 65535}
    for a, b in [(5, 5), (17, 17), (16, 18), (6, 6), (100, 99), (65001, 65535), (4001, 4061), (30001, 30030)]:
        say("range", a, b)
        got = ffxpy.search(dmap, M, a, b)
        first = max(5, a | 1)
        want = [(s, gold[s]) for s in range(first, b + 1, 2) if s in gold]
        assert [(int(r["sum"]), int(r["low"])) for r in got] == want, (a, b)


def _oracle_solve(sums, map_cap):
    p = run_ffref(["--solve", "--map-cap", map_cap], stdin="\n".join(map(str, sums)).encode(), timeout=900)
    return {int(s): int(lo) for s, lo, _ in (l.split() for l in p.stdout.decode().split("\n") if l)}


@pytest.mark.parametrize("map_value", [2 ** 24 - 1, 2 ** 31 - 1])
def test_slots_match_oracle_beyond_goldens(map_value):
    # ranges up to 2^22 + 2^16 (the batched workload's reach) with a map far
    # below ff_seg's: most probes are answered by Miller-Rabin
    rng = random.Random(11 + map_value % 7)
    ranges = [(a, a + 63) for a in (rng.randrange(2 ** 21, 2 ** 22 + 2 ** 16) | 1 for _ in range(6))]
    ranges += [(5, 3001), (2 ** 20 - 63, 2 ** 20 + 63)]
    top = max(b for _, b in ranges)
    say("beyond goldens", map_value, ranges)
    dmap = ffxpy.sieve_map(map_value)
    plan = ffxpy.SearchPlan(top)
    sums = []
    got = {}
    for a, b in ranges:
        slots = ffxpy.search_slots(dmap, map_value, a, b, plan)
        first = max(5, a | 1)
        for i, v in enumerate(slots):
            s = first + 2 * i
            sums.append(s)
            got[s] = int(v) & ffxpy.SLOT_LOW_MASK
    want = _oracle_solve(sums, 2 ** 31 - 1)
    assert got == want


def test_slots_match_oracle_up_to_the_largest_sum():
    # sums near 2^23 (the served workload's top), 2^25 and FFX_MAX_SUM
    # (2^26 - 1) over a 16 MiB map: probe values reach s^2/4 < 2^50, the
    # top of the double-precision Miller-Rabin range
    ranges = [(2 ** 23 - 63, 2 ** 23 - 1), (2 ** 25 - 79, 2 ** 25 - 1), (2 ** 26 - 79, 2 ** 26 - 1)]
    say("largest sums", ranges)
    M = 2 ** 24 - 1
    dmap = ffxpy.sieve_map(M)
    plan = ffxpy.SearchPlan(2 ** 26 - 1)
    got, sums = {}, []
    for a, b in ranges:
        for i, v in enumerate(ffxpy.search_slots(dmap, M, a, b, plan)):
            s = (a | 1) + 2 * i
            sums.append(s)
            got[s] = int(v) & ffxpy.SLOT_LOW_MASK
    want = _oracle_solve(sums, M)
    assert got == want
    assert sum(1 for v in want.values() if v) >= 4  # emitted sums among them


def test_records_entry_point(c1_map):
    dmap, M = c1_map
    lib = ffxpy.load()
    n = (65535 - 5) // 2 + 1
    rec = ffxpy.DeviceBuffer(n * 32)
    cnt = ffxpy.DeviceBuffer(4)
    ffxpy.check(lib.ffx_memset_async(rec.ptr, 0, n * 32, None))
    ffxpy.check(lib.ffx_memset_async(cnt.ptr, 0, 4, None))
    ffxpy.check(lib.ffx_device_sync())
    ffxpy.check(lib.ffx_search_records(0, dmap.ptr, M, 5, 65535, cnt.ptr, rec.ptr), "ffx_search_records")
    r = rec.to_host(n * 32).view(ffxpy.RECORD_DTYPE)
    assert int(cnt.to_host(4, np.uint32)[0]) == 2357
    gold = {s: lo for s, lo, _ in golden_solutions() if s <= 65535}
    solved = r[r["sum"] != 0]
    assert {int(x["sum"]): int(x["low"]) for x in solved} == gold
    idx = np.nonzero(r["sum"])[0]
    assert np.array_equal(r["sum"][idx], 5 + 2 * idx)  # sum-indexed slots


def test_compaction_preserves_slot_order():
    lib = ffxpy.load()
    rng = np.random.default_rng(5)
    for n in [1, 63, 4096, 4097, 100000, 3 * 4096 * 1024 + 17]:
        say("compact", n)
        slots = np.where(rng.random(n) < 0.07, rng.integers(1, 1 << 20, n), 0).astype(np.uint32)
        d = ffxpy.DeviceBuffer(n * 4)
        d.from_host(slots)
        out = ffxpy.DeviceBuffer(n * 16)
        cnt = ffxpy.DeviceBuffer(4)
        plan = ffxpy.SearchPlan(1000)
        ffxpy.check(lib.ffx_compact_async(plan.h, d.ptr, n, 1001, out.ptr, cnt.ptr, None))
        ffxpy.check(lib.ffx_device_sync())
        c = int(cnt.to_host(4, np.uint32)[0])
        idx = np.nonzero(slots)[0]
        assert c == len(idx)
        got = out.to_host(c * 16).view(ffxpy.SOLUTION_DTYPE) if c else np.zeros(0, ffxpy.SOLUTION_DTYPE)
        assert np.array_equal(got["sum"], (1001 + 2 * idx).astype(np.uint32))
        assert np.array_equal(got["low"], slots[idx] & ffxpy.SLOT_LOW_MASK)


def test_search_rejects_bad_arguments(c1_map):
    dmap, M = c1_map
    lib = ffxpy.load()
    plan = ffxpy.SearchPlan(1000)
    out = ffxpy.DeviceBuffer(64)
    assert lib.ffx_search_slots_async(plan.h, dmap.ptr, M, 5, 5001, out.ptr, None) == -34  # above the plan
    assert lib.ffx_search_slots_async(plan.h, dmap.ptr, 7, 5, 99, out.ptr, None) == -22    # map too small


def _listed_splits(s, is_prime):
    """The splits the kernel lists for sum s, enumerated d by d: for each
    k from the largest with 3*2^k <= s-3 down to 2, every odd d in
    [dmin, (s-3) >> k] whose class mod 30 the family and the probe values'
    residues mod 3 and 5 admit (the d = o split's class kept whole).
    None when statement 1 or two Power2 primes end the sum first."""
    if is_prime[s - 2] or (s % 3 == 0 and not is_prime[s // 3]):
        return None
    if sum(1 for k in range(2, 32) if (1 << k) < s - 2 and is_prime[s - (1 << k)]) > 1:
        return None
    total = 0
    k = ((s - 3) // 3).bit_length() - 1
    while k >= 2:
        D = (s - 3) >> k
        dmin = 5 if k & 1 else 3
        d = np.arange(dmin | 1, D + 1, 2, dtype=np.int64)
        o = s - (d << k)
        x, y = (1 << k) + d * o - 2, d + o * (1 << k) - 2
        ok = np.ones(len(d), dtype=bool)
        if k & 1:
            ok &= (d % 3 != 0) & (o % 3 != 0)
        # the rule is per class of d mod 30 (o, x, y mod 15 follow from it);
        # the d = o split's class keeps its y unchecked
        dstar = s // ((1 << k) + 1) if s % ((1 << k) + 1) == 0 else -1
        star_cls = (dstar % 30) if dstar >= dmin and dstar % 2 == 1 and dstar <= D else -1
        ok &= (x % 3 != 0) & (x % 5 != 0) & (((y % 3 != 0) & (y % 5 != 0)) | (d % 30 == star_cls))
        total += int(ok.sum())
        k -= 1
    return total


def test_listed_splits_are_exactly_the_splits_with_d_up_to_its_bound():
    # the count of listed splits per segment is P + (offset <= rem) per class
    # with P = D / 30; counting (D + 1) / 30 periods listed a whole extra
    # period, d beyond (s - 3) >> k, whenever 30 divides D + 1.  The counters
    # build sums the kernel's listed splits over a launch; here they equal
    # the splits enumerated d by d, over sums where 30 | D + 1 for some k.
    import ctypes as C
    import os
    lim = 1 << 17
    sieve = np.ones(lim, dtype=bool)
    sieve[:2] = False
    for p in range(2, int(lim ** 0.5) + 1):
        if sieve[p]:
            sieve[p * p::p] = False
    lo, hi = 32769, 36767  # odd sums from 2^15 + 1: long k-segments, scanned
    want = 0
    hits = 0
    for s in range(lo, hi + 1, 2):
        n = _listed_splits(s, sieve)
        if n is not None:
            want += n
            hits += any(((s - 3) >> k) % 30 == 29 for k in range(2, 16))
    assert hits >= 50  # 53 sums reach the splits with 30 | D + 1 for some k
    lib = C.CDLL(os.path.join(os.path.dirname(ffxpy.LIB_PATH), "libffx_counters.so"))
    lib.ffx_init.argtypes = [C.c_int]
    lib.ffx_probe_counters.argtypes = [C.c_void_p, C.c_int]
    lib.ffx_search_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_uint32]
    lib.ffx_search_plan_destroy.argtypes = [C.c_void_p]
    lib.ffx_search_slots_async.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p,
                                           C.c_void_p]
    assert lib.ffx_init(0) == 0
    M = 268435471
    dmap = ffxpy.sieve_map(M)
    plan = C.c_void_p()
    assert lib.ffx_search_plan_create(C.byref(plan), hi) == 0
    slots = ffxpy.DeviceBuffer(((hi - lo) // 2 + 1) * 4)
    counts = np.zeros(16, dtype=np.uint64)
    assert lib.ffx_probe_counters(counts.ctypes.data, 1) == 0
    say("listed splits of odd sums", lo, hi)
    assert lib.ffx_search_slots_async(plan, dmap.ptr, M, lo, hi, slots.ptr, None) == 0
    assert lib.ffx_device_sync() == 0
    assert lib.ffx_probe_counters(counts.ctypes.data, 1) == 0
    lib.ffx_search_plan_destroy(plan)
    slots.close()
    dmap.close()
    assert int(counts[9]) == want
