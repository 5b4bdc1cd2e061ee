"""The allocation-free search entry point (ffx_search_records_noalloc)
against ffx_search_records through the C ABI: byte-equal records and equal
counts on two calls in a row (the second over a smaller range, so stale
slots of the first call would show), the golden solutions, and the size
limit of its fixed scratch array."""
import sys

import numpy as np
import pytest

import ffxpy
from conftest import golden_solutions

pytestmark = pytest.mark.gpu

MAX_SUM = 0x1FFFFF  # FFX_NOALLOC_MAX_SUM
M_2P21 = 274877906959  # ff_seg geometry of odd sums 5..2097151 (16 GiB map)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def run_records(fn, dmap, M, a, b, stream=None):
    lib = ffxpy.load()
    n = (b - a) // 2 + 1
    rec = ffxpy.DeviceBuffer(n * 32)
    cnt = ffxpy.DeviceBuffer(4)
    ffxpy.check(lib.ffx_memset_async(rec.ptr, 0, n * 32, None))
    ffxpy.check(lib.ffx_memset_async(cnt.ptr, 0, 4, None))
    ffxpy.check(lib.ffx_device_sync())
    if fn == "noalloc":
        ffxpy.check(lib.ffx_search_records_noalloc(0, dmap.ptr, M, a, b, cnt.ptr, rec.ptr, stream),
                    "ffx_search_records_noalloc")
        ffxpy.check(lib.ffx_stream_sync(stream) if stream else lib.ffx_device_sync())
    else:
        ffxpy.check(lib.ffx_search_records(0, dmap.ptr, M, a, b, cnt.ptr, rec.ptr), "ffx_search_records")
    out = rec.to_host(n * 32).view(ffxpy.RECORD_DTYPE).copy(), int(cnt.to_host(4, np.uint32)[0])
    rec.close()
    cnt.close()
    return out


def test_noalloc_matches_records_on_two_calls():
    ffxpy.init(0)
    lib = ffxpy.load()
    say("sieve the 16 GiB map of odd sums 5..2097151")
    dmap = ffxpy.sieve_map(M_2P21)
    stream = lib.ffx_stream_create()
    assert stream
    try:
        # first call: the largest range the scratch array holds
        say("no-alloc records 5..", MAX_SUM)
        r1, c1 = run_records("noalloc", dmap, M_2P21, 5, MAX_SUM, stream)
        say("plan records 5..", MAX_SUM)
        w1, d1 = run_records("plan", dmap, M_2P21, 5, MAX_SUM)
        assert c1 == d1 == 71424  # solution lines of the 2097152 golden leg
        assert r1.tobytes() == w1.tobytes()
        gold = {s: lo for s, lo, _ in golden_solutions()}
        low = r1[(r1["sum"] != 0) & (r1["sum"] <= 1048576)]
        assert {int(x["sum"]): int(x["low"]) for x in low} == gold
        # second call, smaller and starting elsewhere: the array keeps the
        # first call's slots, which must not leak into this one's records
        say("no-alloc records 1001..65535")
        r2, c2 = run_records("noalloc", dmap, M_2P21, 1001, 65535, stream)
        w2, d2 = run_records("plan", dmap, M_2P21, 1001, 65535)
        assert c2 == d2 == sum(1 for s in gold if 1001 <= s <= 65535)
        assert r2.tobytes() == w2.tobytes()
        idx = np.nonzero(r2["sum"])[0]
        assert np.array_equal(r2["sum"][idx], 1001 + 2 * idx)
    finally:
        ffxpy.check(lib.ffx_stream_destroy(stream))
        dmap.close()


def test_noalloc_rejects_sizes_beyond_its_array():
    ffxpy.init(0)
    lib = ffxpy.load()
    dmap = ffxpy.sieve_map(268435471)
    rec = ffxpy.DeviceBuffer(32)
    cnt = ffxpy.DeviceBuffer(4)
    say("no-alloc records beyond the maximum")
    rc = lib.ffx_search_records_noalloc(0, dmap.ptr, 268435471, 5, MAX_SUM + 2, cnt.ptr, rec.ptr, None)
    assert rc == -34  # FFX_ERANGE, before any launch
    assert b"maximum" in lib.ffx_last_error()
    rc = lib.ffx_search_records_noalloc(0, dmap.ptr, 268435471, 6, 101, cnt.ptr, rec.ptr, None)
    assert rc == -22  # even sumStart
    for b in (rec, cnt, dmap):
        b.close()
