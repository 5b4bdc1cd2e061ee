#!/usr/bin/env python3
"""Bench: plain search of odd sums [5, 2^20) on one MI355X.

One step = the hot path over the whole workload with nothing resident but
the plans' small tables: sieve the 4 GiB odd-only prime bitmap (values up to
M = 68,719,476,751, ff_seg's geometry for sumLimit 1,048,575), search all
524,286 odd sums, compact the solutions in sum order.  `value` is odd sums
per second of whole steps.  Kernel times come from HIP events on the stream
the kernels run on.  Parity is checked on every run: the emitted text of the
last step hashes to the reference's golden.

Two more legs follow (each its own ff_sieve process, after the plain
buffers are freed): the batched workload (4096 seeded ranges in one
graph-captured persistent launch) and the served workload (20000 seeded
Poisson requests, p50/p99 latency beside sequential plain launches); their
summaries are attached as "batched" and "served" (the map sieved before the
first arrival, as the plain launches have it resident), plus
"served_on_demand" (the map grown while requests arrive).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sums2p20|sums2p21] [--no-cpu] [--no-modes]
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ff-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "sums searched/s + sieve HBM GB/s; served p50/p99 latency vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (sumLimit, golden leg, description)
    "sums2p20": (1048575, 1048576, "plain search, odd sums [5, 2^20): one sieve pass (4 GiB map) + one search launch"),
    "sums2p21": (2097151, 2097152, "plain search, odd sums [5, 2^21): one sieve pass (16 GiB map) + one search launch"),
}

def load_ref_work():
    p = os.path.join(ROOT, "profiles", "reference_work.json")
    if os.path.exists(p):
        return json.load(open(p))
    return {}


def traffic_for(kernel, workload):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    t = json.load(open(p)).get(workload, {}).get(kernel)
    return t


def probes_for(workload):
    """Map reads of one search launch, by kind (profiles/search_probes.json,
    written by `bench.py --count-probes` with the counters build)."""
    p = os.path.join(ROOT, "profiles", "search_probes.json")
    if not os.path.exists(p):
        return None
    return json.load(open(p)).get(workload)


def random_read_ceiling():
    """The device's rate of random single-byte reads over a 4 GiB buffer
    (bin/ff_random_reads), the access pattern of the search's map probes."""
    exe = os.path.join(ROOT, "ff-gpu_amd", "bin", "ff_random_reads")
    if not os.path.exists(exe):
        return None
    try:
        p = subprocess.run([exe], capture_output=True, timeout=120)
    except subprocess.TimeoutExpired:
        return None
    if p.returncode != 0:
        return None
    return json.loads(p.stdout.decode().strip().splitlines()[-1])


COUNTER_NAMES = ["sums_enumerated", "splits_visited", "splits_past_residue_filter", "reads_2k_plus_do",
                 "reads_d_plus_o2k", "reads_o", "allbutone_checks", "reads_in_allbutone", "miller_rabin_tests",
                 "splits_listed"]


def count_probes(workload):
    """One search of the workload with lib/libffx_counters.so: the counts of
    map reads and the other work by kind (the kernel's own, not the
    reference's order)."""
    import numpy as np
    import ffxpy
    ffxpy.LIB_PATH = os.path.join(ROOT, "ff-gpu_amd", "lib", "libffx_counters.so")
    lib = ffxpy.load()
    lib.ffx_probe_counters.restype = C.c_int
    lib.ffx_probe_counters.argtypes = [C.c_void_p, C.c_int]
    ffxpy.init(0)
    sum_limit = WORKLOADS[workload][0]
    M = 68719476751 if workload == "sums2p20" else 274877906959
    dmap = ffxpy.sieve_map(M)
    arr = np.zeros(16, dtype=np.uint64)
    ffxpy.check(lib.ffx_probe_counters(arr.ctypes.data, 1))
    sol = ffxpy.search(dmap, M, 5, sum_limit)
    ffxpy.check(lib.ffx_probe_counters(arr.ctypes.data, 1))
    res = {n: int(arr[i]) for i, n in enumerate(COUNTER_NAMES)}
    res["map_reads"] = res["reads_2k_plus_do"] + res["reads_d_plus_o2k"] + res["reads_o"] + res["reads_in_allbutone"]
    res["solutions"] = int(len(sol))
    print(json.dumps({workload: res}))


def cpu_baseline(sum_limit, timeout_s):
    """The reference program (oracle/_ref/ff_seg, built from the reference's
    source, 31 threads) on the same workload, timed by its own clocks."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ff_seg")
    kind = "reference"
    if not os.path.exists(exe):
        exe = os.path.join(ROOT, "oracle", "ffref")
        kind = "port"
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        t0 = time.time()
        try:
            p = subprocess.run([exe, "5", str(sum_limit)], cwd=d, capture_output=True, timeout=timeout_s)
        except subprocess.TimeoutExpired:
            return {"value": None, "unit": "sums/s", "cores": 31, "kind": kind,
                    "sample": f"timed out after {timeout_s} s"}
        wall = time.time() - t0
    out = p.stdout.decode(errors="replace")
    pt = int(re.search(r"Prime time: (\d+)", out).group(1))
    ft = int(re.search(r"Freudenthal time: (\d+)", out).group(1))
    nodd = (sum_limit - 5) // 2 + 1
    mapbytes = (2 ** 36 + 16) // 16 if sum_limit == 1048575 else None
    return {
        "value": nodd / ((pt + ft) * 1e-6),
        "unit": "sums/s",
        "cores": 31 if kind == "reference" else os.cpu_count(),
        "kind": kind,
        "sample": f"full workload 5..{sum_limit} ({nodd} odd sums), one run; ff_seg's own Prime/Freudenthal "
                  f"clocks: sieve {pt / 1e6:.3f} s, search {ft / 1e6:.3f} s, wall {wall:.2f} s; host has "
                  f"{os.cpu_count()} CPUs",
        "sieve_s": pt / 1e6,
        "search_s": ft / 1e6,
        "search_sums_per_s": nodd / (ft * 1e-6),
        "host_cpus": os.cpu_count(),
    }


def mode_leg(args, timeout_s):
    """One batched or served run of ff_sieve (its own process, its own
    device buffers); returns its JSON summary line."""
    exe = os.path.join(ROOT, "ff-gpu_amd", "bin", "ff_sieve")
    try:
        p = subprocess.run([exe] + args, capture_output=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s} s"}
    if p.returncode != 0:
        return {"error": p.stderr.decode(errors="replace")[-500:]}
    return json.loads(p.stdout.decode().strip().splitlines()[-1])


def calibrate_copy(nbytes):
    """A plain device-to-device copy of known size (torch's copy kernel),
    three times: FETCH_SIZE and WRITE_SIZE of one dispatch against nbytes
    give the counters' scale."""
    import torch
    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": nbytes, "kib": nbytes / 1024}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sums2p20", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-timeout", type=float, default=150.0)
    ap.add_argument("--no-modes", action="store_true", help="skip the batched and served legs")
    ap.add_argument("--count-probes", action="store_true",
                    help="only count the search's map reads by kind (needs make -C ff-gpu_amd counters)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only copy a BYTES device buffer three times: the run that scales the HBM byte counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.count_probes:
        count_probes(args.workload)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the sieve and search are single-device here (each GPU would "
              "re-sieve locally and search a disjoint sum range); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ffxpy
    from conftest import golden_sha  # noqa: E402  (the golden hash table only)

    sum_limit, leg, desc = WORKLOADS[args.workload]
    lib = ffxpy.load()
    ffxpy.init(0)
    dev = ffxpy.device_info(0)
    M = 68719476751 if args.workload == "sums2p20" else 274877906959
    map_bytes = (M >> 4) + 1
    nodd = (sum_limit - 5) // 2 + 1

    sp = ffxpy.SievePlan(M)
    qp = ffxpy.SearchPlan(sum_limit)
    dmap = ffxpy.DeviceBuffer((map_bytes + 255) & ~255)
    slots = ffxpy.DeviceBuffer(nodd * 4)
    out = ffxpy.DeviceBuffer(nodd * 16)
    cnt = ffxpy.DeviceBuffer(4)
    stream = lib.ffx_stream_create()
    ev = [lib.ffx_event_create() for _ in range(4)]

    def step():
        ffxpy.check(lib.ffx_event_record(ev[0], stream))
        ffxpy.check(lib.ffx_sieve_map_async(sp.h, dmap.ptr, map_bytes, stream), "sieve")
        ffxpy.check(lib.ffx_event_record(ev[1], stream))
        ffxpy.check(lib.ffx_search_slots_async(qp.h, dmap.ptr, M, 5, sum_limit, slots.ptr, stream), "search")
        ffxpy.check(lib.ffx_event_record(ev[2], stream))
        ffxpy.check(lib.ffx_compact_async(qp.h, slots.ptr, nodd, 5, out.ptr, cnt.ptr, stream), "compact")
        ffxpy.check(lib.ffx_event_record(ev[3], stream))

    def elapsed(a, b):
        ms = C.c_double()
        ffxpy.check(lib.ffx_event_elapsed_ms(ev[a], ev[b], C.byref(ms)))
        return ms.value

    for _ in range(args.warmup):
        step()
    ffxpy.check(lib.ffx_stream_sync(stream))

    sieve_ms, search_ms, compact_ms = [], [], []
    ffxpy.check(lib.ffx_device_sync())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        ffxpy.check(lib.ffx_event_sync(ev[3]))  # per-step events are reused; read them before the next step
        sieve_ms.append(elapsed(0, 1))
        search_ms.append(elapsed(1, 2))
        compact_ms.append(elapsed(2, 3))
    ffxpy.check(lib.ffx_device_sync())
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps

    # parity of the last step: rebuild ff_seg's stdout from the solutions
    c = int(cnt.to_host(4, np.uint32)[0])
    sol = out.to_host(c * 16).view(ffxpy.SOLUTION_DTYPE)
    seg = subprocess.run([os.path.join(ROOT, "ff-gpu_amd", "bin", "ff_hostcheck"), "seg", "5", str(sum_limit)],
                         capture_output=True, text=True).stdout
    headers = [l[7:] for l in seg.splitlines() if l.startswith("header ")]
    text = "".join(h + "\n" for h in headers)
    text += "".join(ffxpy.format_line(i + 1, int(r["sum"]), int(r["low"]), int(r["high"]), int(r["tags"]))
                    for i, r in enumerate(sol))
    text += "Prime time: N μs\nFreudenthal time: N μs\n"
    parity = hashlib.sha256(text.encode()).hexdigest() == golden_sha(leg)

    sieve_avg = sum(sieve_ms) / len(sieve_ms)
    search_avg = sum(search_ms) / len(search_ms)
    work = load_ref_work().get(args.workload, {})
    probes = work.get("probes")
    # dominant kernel: the search; its algorithmic bytes are the reference's
    # bitmap probes at one 64-byte line each
    if probes:
        ach = probes * 64 / (search_avg * 1e-3) / 1e9
        traffic = traffic_for("search_sums", args.workload)
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "search_sums", "algorithmic": f"{probes} reference-order bitmap probes x 64 B",
                "traffic_gbs": round(traffic / (search_avg * 1e-3) / 1e9, 1) if traffic else None,
                "note": "frac above 1: the kernel's residue filter and family-aware split lists skip most of the "
                        "reference's probes, so it reads far fewer bytes (traffic, measured by rocprofv3 "
                        "FETCH_SIZE/WRITE_SIZE) than the reference-order work implies; its map reads run near "
                        "the device's random-read ceiling (random_read_roof) with VALU issue near its peak "
                        "(VALU wave-instructions per CU per clock in profiles/pmc_plain_2p20.md): both bind"}
    else:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": "search_sums", "algorithmic": "reference probe count missing"}
    sieve_gbs = map_bytes / (sieve_avg * 1e-3) / 1e9
    roof_sieve = {"bound": "hbm", "achieved": round(sieve_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(sieve_gbs / HBM_PEAK_GBS, 4), "traffic": traffic_for("sieve_segments", args.workload),
                  "kernel": "sieve_segments", "algorithmic": f"{map_bytes} map bytes written once",
                  "note": "HBM is not what binds the sieve: its strikes are LDS atomics and its issue runs near "
                          "the VALU peak (VALU wave-instructions per CU per clock in profiles/pmc_plain_2p20.md)"}

    for b in (dmap, slots, out, cnt):
        b.close()
    sp.close()
    qp.close()
    # the search's map reads (counted) against the device's measured rate of
    # random single-byte reads: the roof for this access pattern, where every
    # read costs a whole 128-byte line
    ceiling = random_read_ceiling()
    pc = probes_for(args.workload)
    if ceiling and pc:
        rate = pc["map_reads"] / (search_avg * 1e-3)
        roof["random_read_roof"] = {"map_reads_per_launch": pc["map_reads"], "achieved_reads_per_s": round(rate, -6),
                                    "ceiling_reads_per_s": round(ceiling["reads_per_s"], -6),
                                    "frac": round(rate / ceiling["reads_per_s"], 3),
                                    "ceiling_from": "bin/ff_random_reads (4 GiB buffer, 2^28 reads, best of 3)"}
    cpu = None if args.no_cpu else cpu_baseline(sum_limit, args.cpu_timeout)
    batched = served = served_on_demand = None
    if not args.no_modes:
        # 4096 seeded ranges (starts below 2^22, widths 2^8..2^16) in one
        # graph-captured persistent launch; the served Poisson stream of
        # 20000 requests (sumLimit up to 2^23, widths 2^6..2^18, 2000/s)
        # with 100 of them replayed as sequential plain launches; in served
        # the map is sieved whole before the first arrival (as the plain
        # launches have it resident), in served_on_demand it is grown while
        # requests arrive
        batched = mode_leg(["--batch", "--opt", "reps=3"], 300)
        served = mode_leg(["--serve", "--opt", "warm=1", "--opt", "plain_sample=100"], 300)
        served_on_demand = mode_leg(["--serve"], 300)
    res = {
        "metric": METRIC,
        "value": nodd / (ms_per_step * 1e-3),
        "unit": "sums/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: the workload is the sum range itself (no input data)",
        "config": {"workload": desc, "sumStart": 5, "sumLimit": sum_limit, "odd_sums": nodd,
                   "maxPrimeMapValue": M, "map_bytes": map_bytes, "device": dev.name.decode() or dev.arch.decode(),
                   "cus": dev.multiProcessorCount},
        "roofline": roof,
        "roofline_sieve": roof_sieve,
        "sieve_ms": round(sieve_avg, 4),
        "sieve_gbs": round(sieve_gbs, 1),
        "search_ms": round(search_avg, 4),
        "search_sums_per_s": nodd / (search_avg * 1e-3),
        "compact_ms": round(sum(compact_ms) / len(compact_ms), 4),
        "solutions": c,
        "parity": "sha256 of emitted stdout == reference golden" if parity else "MISMATCH",
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "served_on_demand": served_on_demand,
    }
    if cpu and cpu.get("value"):
        res["gpu_over_cpu"] = round(res["value"] / cpu["value"], 1)
    print(json.dumps(res))
    if not parity:
        sys.exit(1)


if __name__ == "__main__":
    main()
