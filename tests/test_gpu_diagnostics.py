"""The diagnostic builds (GPU): the counters library answers exactly as libffx
and its counts are consistent; the random-read ceiling program runs.

The counters library exports the same entry points as libffx, so it runs
in a child process of its own (one libffx per process)."""
import json
import os
import subprocess
import sys

import pytest

from conftest import PKG, ROOT

pytestmark = pytest.mark.gpu

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import ffxpy
ffxpy.LIB_PATH = sys.argv[2]
lib = ffxpy.load()
import ctypes as C
lib.ffx_probe_counters.restype = C.c_int
lib.ffx_probe_counters.argtypes = [C.c_void_p, C.c_int]
ffxpy.init(0)
M = 268435471  # ff_seg's map bound for odd sums 5..65535
dmap = ffxpy.sieve_map(M)
arr = np.zeros(16, dtype=np.uint64)
assert lib.ffx_probe_counters(arr.ctypes.data, 1) == 0
sol = ffxpy.search(dmap, M, 5, 65535)
assert lib.ffx_probe_counters(arr.ctypes.data, 1) == 0
again = np.zeros(16, dtype=np.uint64)
assert lib.ffx_probe_counters(again.ctypes.data, 0) == 0
print(json.dumps({"counts": [int(x) for x in arr], "after_reset": [int(x) for x in again],
                  "solutions": [[int(r["sum"]), int(r["low"])] for r in sol]}))
"""


def test_counters_build_matches_and_counts():
    lib = os.path.join(PKG, "lib", "libffx_counters.so")
    if not os.path.exists(lib):
        pytest.skip("lib/libffx_counters.so not built (make -C ff-gpu_amd counters)")
    print("[case] counters build, odd sums 5..65535", flush=True)
    p = subprocess.run([sys.executable, "-c", CHILD, PKG, lib], capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    r = json.loads(p.stdout.strip().splitlines()[-1])
    c = r["counts"]
    import ffxpy
    ffxpy.init(0)
    M = 268435471
    want = [[int(x["sum"]), int(x["low"])] for x in ffxpy.search(ffxpy.sieve_map(M), M, 5, 65535)]
    assert r["solutions"] == want  # the counters build answers as libffx does
    sums, visited, survivors, rx, ry, ro, ab1, rab1, mr, listed = c[:10]
    assert 0 < sums <= 32766 and 0 < visited <= listed
    assert survivors <= visited and rx <= survivors  # every first read comes from a filter survivor
    assert ry <= rx and ab1 <= rx and rab1 > 0
    assert r["after_reset"] == [0] * 16


def test_random_read_ceiling_runs():
    exe = os.path.join(PKG, "bin", "ff_random_reads")
    print("[case] ff_random_reads 64 MiB", flush=True)
    p = subprocess.run([exe, str(64 << 20)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    r = json.loads(p.stdout.strip().splitlines()[-1])
    assert r["buffer_bytes"] == 64 << 20 and r["reads_per_launch"] == 2 ** 28
    assert 1e9 < r["reads_per_s"] < 1e13
