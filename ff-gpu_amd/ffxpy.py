"""ctypes binding of libffx (include/ffx.h) for the tests and the bench.

The product is the C ABI in lib/libffx.so; this module only declares its
signatures and wraps the return-code convention (0 = ok, negative errno +
ffx_last_error()).  There is no fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libffx.so")
FF_SIEVE = os.path.join(HERE, "bin", "ff_sieve")
FF_HOSTCHECK = os.path.join(HERE, "bin", "ff_hostcheck")
HEADER = os.path.join(ROOT, "include", "ffx.h")

TAG_LOW_PRIME = 0x10000000
TAG_LOW_POW2 = 0x20000000
TAG_HIGH_PRIME = 0x40000000
TAG_HIGH_POW2 = 0x80000000
SLOT_LOW_MASK = 0x0FFFFFFF

SOLUTION_DTYPE = np.dtype([("sum", "<u4"), ("low", "<u4"), ("high", "<u4"), ("tags", "<u4")])
RECORD_DTYPE = np.dtype([("sum", "<u4"), ("pad0", "<u4"), ("low", "<u8"), ("high", "<u8"), ("tag", "u1"),
                         ("pad1", "u1", (7,))])


class FfxError(RuntimeError):
    pass


class DeviceInfo(C.Structure):
    _fields_ = [("vendor", C.c_char * 8), ("name", C.c_char * 256), ("arch", C.c_char * 64),
                ("busId", C.c_char * 32), ("freeBytes", C.c_uint64), ("totalBytes", C.c_uint64),
                ("sharedMemPerBlock", C.c_uint64), ("sharedMemPerMultiprocessor", C.c_uint64),
                ("l2Bytes", C.c_uint64), ("multiProcessorCount", C.c_int32), ("maxThreadsPerBlock", C.c_int32),
                ("warpSize", C.c_int32), ("clockKHz", C.c_int32), ("memoryClockKHz", C.c_int32),
                ("memoryBusWidth", C.c_int32), ("runtimeIndex", C.c_int32), ("reserved", C.c_int32)]


class Interval(C.Structure):
    _fields_ = [("first", C.c_uint32), ("last", C.c_uint32), ("slotOffset", C.c_uint64)]


_vp, _u32, _u64, _i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
_SIGS = {
    "ffx_abi_version": (_i32, []),
    "ffx_init": (_i32, [_i32]),
    "ffx_last_error": (C.c_char_p, []),
    "ffx_device_count": (_i32, [C.POINTER(_i32)]),
    "ffx_device_info_get": (_i32, [_i32, C.POINTER(DeviceInfo)]),
    "ffx_mem_info": (_i32, [C.POINTER(_u64), C.POINTER(_u64)]),
    "ffx_alloc": (_vp, [_u64]),
    "ffx_free": (_i32, [_vp]),
    "ffx_memset_async": (_i32, [_vp, _i32, _u64, _vp]),
    "ffx_copy_h2d": (_i32, [_vp, _vp, _u64]),
    "ffx_copy_d2h": (_i32, [_vp, _vp, _u64]),
    "ffx_copy_h2d_async": (_i32, [_vp, _vp, _u64, _vp]),
    "ffx_copy_d2h_async": (_i32, [_vp, _vp, _u64, _vp]),
    "ffx_alloc_pinned": (_vp, [_u64]),
    "ffx_free_pinned": (_i32, [_vp]),
    "ffx_stream_create": (_vp, []),
    "ffx_stream_destroy": (_i32, [_vp]),
    "ffx_stream_sync": (_i32, [_vp]),
    "ffx_event_create": (_vp, []),
    "ffx_event_destroy": (_i32, [_vp]),
    "ffx_event_record": (_i32, [_vp, _vp]),
    "ffx_event_wait": (_i32, [_vp, _vp]),
    "ffx_event_sync": (_i32, [_vp]),
    "ffx_event_query": (_i32, [_vp]),
    "ffx_event_elapsed_ms": (_i32, [_vp, _vp, C.POINTER(C.c_double)]),
    "ffx_device_sync": (_i32, []),
    "ffx_sieve_plan_create": (_i32, [C.POINTER(_vp), _u64]),
    "ffx_sieve_plan_create_from_primes": (_i32, [C.POINTER(_vp), C.POINTER(_u32), _u32, _u64]),
    "ffx_sieve_plan_destroy": (_i32, [_vp]),
    "ffx_sieve_slab_async": (_i32, [_vp, _u64, _u64, _vp, _vp]),
    "ffx_sieve_map_async": (_i32, [_vp, _vp, _u64, _vp]),
    "ffx_sieve_map_range_async": (_i32, [_vp, _vp, _u64, _u64, _vp]),
    "ffx_search_plan_create": (_i32, [C.POINTER(_vp), _u32]),
    "ffx_search_plan_destroy": (_i32, [_vp]),
    "ffx_search_slots_async": (_i32, [_vp, _vp, _u64, _u32, _u32, _vp, _vp]),
    "ffx_compact_async": (_i32, [_vp, _vp, _u64, _u32, _vp, _vp, _vp]),
    "ffx_search_async": (_i32, [_vp, _vp, _u64, _u32, _u32, _vp, _vp, _vp]),
    "ffx_search_records": (_i32, [_i32, _vp, _u64, _u64, _u64, _vp, _vp]),
    "ffx_search_records_noalloc": (_i32, [_i32, _vp, _u64, _u32, _u32, _vp, _vp, _vp]),
    "ffx_search_intervals_async": (_i32, [_vp, _vp, _u64, _vp, _u32, _u64, _u32, _vp, _vp]),
    "ffx_search_intervals_by_sum_async": (_i32, [_vp, _vp, _u64, _vp, _u32, _u64, _u32, _u32, _vp, _u64, _vp]),
    "ffx_search_plan_set": (_i32, [_vp, _i32, _u32]),
    "ffx_search_plan_get": (_i32, [_vp, _i32, C.POINTER(_u32)]),
    "ffx_sieve_plan_set": (_i32, [_vp, _i32, _u32]),
    "ffx_sieve_plan_get": (_i32, [_vp, _i32, C.POINTER(_u32)]),
    "ffx_graph_begin": (_i32, [_vp]),
    "ffx_graph_end": (_i32, [_vp, C.POINTER(_vp)]),
    "ffx_graph_launch": (_i32, [_vp, _vp]),
    "ffx_graph_destroy": (_i32, [_vp]),
}

SEARCH_GRAB, SEARCH_EAGER_SPLITS, SEARCH_BLOCKS_PER_CU = 1, 2, 3
SIEVE_SEGMENT_BYTES, SIEVE_BLOCKS_PER_CU, SIEVE_LANE_CARRY, SIEVE_HOLD_PRIMES = 1, 2, 3, 4

_lib = None


def header_symbols() -> list[str]:
    """Every function name include/ffx.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ffx_[a-z0-9_]+)\s*\(", text)))


def load():
    """Loads lib/libffx.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C ff-gpu_amd)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().ffx_last_error().decode(errors="replace")
        raise FfxError(f"{what or 'ffx'} failed ({rc}): {msg}")


def ptr(x) -> int:
    return x.ptr if isinstance(x, DeviceBuffer) else int(x)


class DeviceBuffer:
    """Device allocation from ffx_alloc, freed on close()/GC."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = load().ffx_alloc(max(1, self.nbytes))
        if not p:
            raise FfxError(f"ffx_alloc({nbytes}) failed: {load().ffx_last_error().decode()}")
        self.ptr = p

    def close(self):
        if self.ptr:
            load().ffx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def to_host(self, nbytes: int | None = None, dtype=np.uint8, offset: int = 0) -> np.ndarray:
        n = self.nbytes - offset if nbytes is None else int(nbytes)
        out = np.empty(n, dtype=np.uint8)
        check(load().ffx_copy_d2h(out.ctypes.data, self.ptr + offset, n), "ffx_copy_d2h")
        return out.view(dtype)

    def from_host(self, arr: np.ndarray, offset: int = 0):
        arr = np.ascontiguousarray(arr)
        check(load().ffx_copy_h2d(self.ptr + offset, arr.ctypes.data, arr.nbytes), "ffx_copy_h2d")


def init(device: int = 0):
    check(load().ffx_init(device), "ffx_init")


def device_count() -> int:
    n = C.c_int(0)
    check(load().ffx_device_count(C.byref(n)), "ffx_device_count")
    return n.value


def device_info(device: int = 0) -> DeviceInfo:
    d = DeviceInfo()
    check(load().ffx_device_info_get(device, C.byref(d)), "ffx_device_info_get")
    return d


class SievePlan:
    def __init__(self, max_prime_map_value: int, primes: np.ndarray | None = None):
        self.M = int(max_prime_map_value)
        h = C.c_void_p()
        if primes is None:
            check(load().ffx_sieve_plan_create(C.byref(h), self.M), "ffx_sieve_plan_create")
        else:
            p = np.ascontiguousarray(primes, dtype=np.uint32)
            check(load().ffx_sieve_plan_create_from_primes(C.byref(h), p.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                           len(p), self.M), "ffx_sieve_plan_create_from_primes")
        self.h = h.value

    def close(self):
        if self.h:
            load().ffx_sieve_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SearchPlan:
    def __init__(self, sum_limit_max: int):
        h = C.c_void_p()
        check(load().ffx_search_plan_create(C.byref(h), int(sum_limit_max)), "ffx_search_plan_create")
        self.h = h.value

    def close(self):
        if self.h:
            load().ffx_search_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def map_bytes_for(M: int) -> int:
    return (int(M) >> 4) + 1


def sieve_map(M: int, plan: SievePlan | None = None) -> DeviceBuffer:
    """Device bitmap for values 0..M (bytes (M>>4)+1, padded to 256)."""
    plan = plan or SievePlan(M)
    nb = map_bytes_for(M)
    buf = DeviceBuffer((nb + 255) & ~255)
    check(load().ffx_sieve_map_async(plan.h, buf.ptr, nb, None), "ffx_sieve_map_async")
    check(load().ffx_device_sync(), "ffx_device_sync")
    buf.map_bytes = nb
    return buf


def search(dmap, M: int, sum_start: int, sum_limit: int, plan: SearchPlan | None = None) -> np.ndarray:
    """Solutions (structured array, ascending sum) for odd sums in [sum_start, sum_limit]."""
    first = max(5, sum_start | 1)
    n = (sum_limit - first) // 2 + 1 if sum_limit >= first else 0
    plan = plan or SearchPlan(max(sum_limit, 5))
    out = DeviceBuffer(max(1, n) * 16)
    cnt = DeviceBuffer(4)
    check(load().ffx_search_async(plan.h, ptr(dmap), int(M), int(sum_start), int(sum_limit), out.ptr, cnt.ptr, None),
          "ffx_search_async")
    check(load().ffx_device_sync(), "ffx_device_sync")
    c = int(cnt.to_host(4, np.uint32)[0])
    if c == 0:
        return np.zeros(0, dtype=SOLUTION_DTYPE)
    return out.to_host(c * 16).view(SOLUTION_DTYPE).copy()


def search_slots(dmap, M: int, sum_start: int, sum_limit: int, plan: SearchPlan | None = None) -> np.ndarray:
    first = max(5, sum_start | 1)
    n = (sum_limit - first) // 2 + 1
    plan = plan or SearchPlan(max(sum_limit, 5))
    slots = DeviceBuffer(n * 4)
    check(load().ffx_search_slots_async(plan.h, ptr(dmap), int(M), int(sum_start), int(sum_limit), slots.ptr, None),
          "ffx_search_slots_async")
    check(load().ffx_device_sync(), "ffx_device_sync")
    return slots.to_host(n * 4, np.uint32)


def format_line(ordinal: int, s: int, low: int, high: int, tags: int) -> str:
    """ff_seg line format (reference/segmentedSieve.C:816-819, :663-677)."""
    if tags & TAG_LOW_PRIME:
        tl = " (prime), "
    elif tags & TAG_LOW_POW2:
        k = low.bit_length() - 1
        tl = f" (2^{k})," + (" " * (3 if k < 10 else 2))
    else:
        tl = ",         "
    if tags & TAG_HIGH_PRIME:
        th = " (prime)"
    elif tags & TAG_HIGH_POW2:
        th = f" (2^{high.bit_length() - 1})"
    else:
        th = ""
    return (f"{ordinal:7d}) sum ={s:9d}, product ={low * high:16d},  low term ={low:9d}{tl}"
            f"high term ={high:9d}{th}\n")
