"""The reference-side adapter (ff-gpu_amd/bindings/reference_adapter.cpp) on
the GPU: the reference's SievePoolOps table, SieveSlabEngineRunFn and
SearchKernelRunFn as its callers would use them (source/pull_scheduler.cpp
slab loop, source/m4 search launch), answering like libffx and the oracle.
The adapter is built against the reference's headers, so the test is skipped
where it was not built."""
import ctypes as C
import os
import sys
import tempfile

import numpy as np
import pytest

import ffxpy
from conftest import golden_solutions, run_ffref

pytestmark = pytest.mark.gpu

ADAPTER = os.path.join(os.path.dirname(ffxpy.FF_SIEVE), "..", "lib", "libffx_refadapter.so")

_vp, _u32, _u64, _i = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
# SievePoolOps (source/sieve_slab_engine.h:49-98), field order
OPS = [
    ("alloc", C.CFUNCTYPE(_vp, _i, _u64)),
    ("memset", C.CFUNCTYPE(_i, _i, _vp, _i, _u64)),
    ("free", C.CFUNCTYPE(_i, _i, _vp)),
    ("copyH2D", C.CFUNCTYPE(_i, _i, _vp, _vp, _u64)),
    ("copyD2H", C.CFUNCTYPE(_i, _i, _vp, _vp, _u64)),
    ("slabCompute", C.CFUNCTYPE(_i, _i, _vp, _u32, _u64, _u64, _vp, _u64, _i)),
    ("createStream", C.CFUNCTYPE(_vp, _i)),
    ("destroyStream", C.CFUNCTYPE(_i, _i, _vp)),
    ("createEvent", C.CFUNCTYPE(_vp, _i)),
    ("destroyEvent", C.CFUNCTYPE(_i, _i, _vp)),
    ("recordEvent", C.CFUNCTYPE(_i, _i, _vp, _vp)),
    ("waitEvent", C.CFUNCTYPE(_i, _i, _vp, _vp)),
    ("syncEvent", C.CFUNCTYPE(_i, _i, _vp)),
    ("elapsedEvents", C.CFUNCTYPE(_i, _i, _vp, _vp, C.POINTER(C.c_double))),
    ("allocPinned", C.CFUNCTYPE(_i, _i, C.POINTER(_vp), _u64)),
    ("freePinned", C.CFUNCTYPE(_i, _i, _vp)),
    ("copyH2DAsync", C.CFUNCTYPE(_i, _i, _vp, _vp, _u64, _vp)),
    ("copyD2HAsync", C.CFUNCTYPE(_i, _i, _vp, _vp, _u64, _vp)),
    ("slabComputeAsync", C.CFUNCTYPE(_i, _i, _vp, _u32, _u64, _u64, _vp, _u64, _vp, _vp)),
]


class SievePoolOps(C.Structure):
    _fields_ = OPS


@pytest.fixture(scope="module")
def adapter():
    if not os.path.exists(ADAPTER):
        pytest.skip("reference adapter not built (the reference's headers were absent at build time)")
    lib = C.CDLL(ADAPTER)
    lib.SievePoolGet_gfx950.restype = C.POINTER(SievePoolOps)
    lib.SievePoolGet_gfx950.argtypes = [_i]
    lib.SieveSlabEngineGetLaunchFn_gfx950.restype = _vp
    lib.SearchKernelGetLaunchFn_gfx950.restype = _vp
    ffxpy.init(0)
    return lib


def primes_upto(n):
    s = np.ones(n + 1, dtype=bool)
    s[:2] = False
    for p in range(2, int(n ** 0.5) + 1):
        if s[p]:
            s[p * p::p] = False
    return np.nonzero(s)[0].astype(np.uint32)


def oracle_map(sum_limit):
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "m.bin")
        assert run_ffref(["--dump-map", f, 5, sum_limit]).returncode == 0
        return np.fromfile(f, dtype=np.uint8)


def test_pool_ops_slab_loop_matches_oracle(adapter):
    # the pull scheduler's slab loop over the reference's table: pinned
    # staging, async slab compute ordered by events, async D2H
    print("[case] adapter slab loop", flush=True)
    sys.stdout.flush()
    ref = oracle_map(65535)
    M = len(ref) * 16 - 1
    primes = primes_upto(int(M ** 0.5))  # with 2, as the reference passes it
    ops = adapter.SievePoolGet_gfx950(0).contents
    slab_bytes = 1 << 20
    d = ops.alloc(0, slab_bytes)
    assert d
    h = _vp()
    assert ops.allocPinned(0, C.byref(h), slab_bytes) == 0
    s_comp, s_copy = ops.createStream(0), ops.createStream(0)
    ev = ops.createEvent(0)
    out = np.zeros(len(ref), dtype=np.uint8)
    pp = primes.ctypes.data_as(_vp)
    for lo in range(0, len(ref), slab_bytes):
        nb = min(slab_bytes, len(ref) - lo)
        seg_lo, seg_hi = lo * 16, min((lo + nb) * 16, M + 1)
        assert ops.slabComputeAsync(0, pp, len(primes), seg_lo, seg_hi, d, slab_bytes, s_comp, None) == 0
        assert ops.recordEvent(0, ev, s_comp) == 0
        assert ops.waitEvent(0, ev, s_copy) == 0
        assert ops.copyD2HAsync(0, h, d, nb, s_copy) == 0
        assert ops.syncEvent(0, ev) == 0
        lib = ffxpy.load()
        ffxpy.check(lib.ffx_stream_sync(s_copy))
        out[lo:lo + nb] = np.ctypeslib.as_array(C.cast(h, C.POINTER(C.c_uint8)), shape=(nb,))
    assert np.array_equal(out, ref)
    # the synchronous slab entry (primeSlab set: priming is in the kernel)
    assert ops.slabCompute(0, pp, len(primes), 0, 1 << 16, d, slab_bytes, 1) == 0
    got = np.zeros(1 << 12, dtype=np.uint8)
    assert ops.copyD2H(0, got.ctypes.data_as(_vp), d, len(got)) == 0
    assert np.array_equal(got, ref[: 1 << 12])
    for x in (ev,):
        assert ops.destroyEvent(0, x) == 0
    assert ops.destroyStream(0, s_comp) == 0 and ops.destroyStream(0, s_copy) == 0
    assert ops.freePinned(0, h) == 0 and ops.free(0, d) == 0


def test_engine_run_and_search_launch_match_goldens(adapter):
    print("[case] adapter engine + search", flush=True)
    sys.stdout.flush()
    M = 268435471  # ff_seg geometry of 5..65535
    primes = primes_upto(int(M ** 0.5))
    run = C.CFUNCTYPE(_i, _i, _vp, _u32, _u64, _vp)(adapter.SieveSlabEngineGetLaunchFn_gfx950(0))
    host = np.zeros((M >> 4) + 1, dtype=np.uint8)
    assert run(0, primes.ctypes.data_as(_vp), len(primes), M, host.ctypes.data_as(_vp)) == 0
    assert np.array_equal(host, oracle_map(65535))
    dmap = ffxpy.DeviceBuffer(len(host))
    lib = ffxpy.load()
    ffxpy.check(lib.ffx_copy_h2d(dmap.ptr, host.ctypes.data_as(_vp), len(host)))
    n = (65535 - 5) // 2 + 1
    rec = ffxpy.DeviceBuffer(n * 32)
    cnt = ffxpy.DeviceBuffer(4)
    ffxpy.check(lib.ffx_memset_async(rec.ptr, 0, n * 32, None))
    ffxpy.check(lib.ffx_memset_async(cnt.ptr, 0, 4, None))
    ffxpy.check(lib.ffx_device_sync())
    search = C.CFUNCTYPE(_i, _i, _vp, _u64, _u64, _u64, _vp, _vp)(adapter.SearchKernelGetLaunchFn_gfx950(0))
    assert search(0, dmap.ptr, M, 5, 65535, cnt.ptr, rec.ptr) == 0
    r = rec.to_host(n * 32).view(ffxpy.RECORD_DTYPE)
    gold = {s: lo for s, lo, _ in golden_solutions() if s <= 65535}
    assert int(cnt.to_host(4, np.uint32)[0]) == len(gold) == 2357
    assert {int(x["sum"]): int(x["low"]) for x in r[r["sum"] != 0]} == gold
    assert not adapter.SievePoolGet_gfx950(99)  # no such device: NULL, as the reference's getters
