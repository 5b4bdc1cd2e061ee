#!/usr/bin/env python3
"""Bench: plain search of odd sums [5, 2^21) on one MI355X.

One step = the hot path over the whole workload with nothing resident but
the plans' small tables: sieve the 16 GiB odd-only prime bitmap (values up to
M = 274,877,906,959, ff_seg's geometry for sumLimit 2,097,151), search all
1,048,574 odd sums, compact the solutions in sum order.  `value` is odd sums
per second of whole steps.  Kernel times come from HIP events on the stream
the kernels run on.  Parity is checked on every run: the emitted text of the
last step hashes to the reference's golden.

Legs after the plain steps (each its own ff_sieve process, after the plain
buffers are freed, each with its output checked range by range against one
plain slot array: a mismatch makes the bench exit non-zero):
  batched           4096 seeded ranges in one graph-captured persistent launch
  served            20000 seeded Poisson requests at 2000/s, map sieved before
                    the first arrival, 100 requests spread over the stream
                    replayed as sequential plain launches beside them
  served_on_demand  the same stream, map grown while requests arrive
  served_capacity   the same requests at 12000/s: the device is the bound

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sums2p21|sums2p20] [--no-cpu] [--no-modes]
  python bench.py --cpu-full      the reference's full CPU runs (ff_seg, 31 and 16 threads, median of 3)
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import platform
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
LINE_BYTES = 128  # HBM bytes one random single-byte read moves (bin/ff_random_reads + FETCH_SIZE)

WORKLOADS = {
    # name: (sumLimit, M, golden leg, description)
    "sums2p21": (2097151, 274877906959, 2097152,
                 "plain search, odd sums [5, 2^21): one sieve pass (16 GiB map) + one search launch"),
    "sums2p20": (1048575, 68719476751, 1048576,
                 "plain search, odd sums [5, 2^20): one sieve pass (4 GiB map) + one search launch"),
}


def profile_json(name):
    p = os.path.join(ROOT, "profiles", name)
    return json.load(open(p)) if os.path.exists(p) else {}


def random_read_ceiling(nbytes):
    """The device's rate of random single-byte reads over a buffer of the
    map's size (bin/ff_random_reads), the access pattern of the search's map
    reads."""
    exe = os.path.join(ROOT, "ff-gpu_amd", "bin", "ff_random_reads")
    if not os.path.exists(exe):
        return None
    try:
        p = subprocess.run([exe, str(nbytes)], capture_output=True, timeout=120)
    except subprocess.TimeoutExpired:
        return None
    if p.returncode != 0:
        return None
    return json.loads(p.stdout.decode().strip().splitlines()[-1])


COUNTER_NAMES = ["sums_enumerated", "splits_visited", "splits_past_residue_filter", "reads_2k_plus_do",
                 "reads_d_plus_o2k", "reads_o", "allbutone_checks", "reads_in_allbutone", "miller_rabin_tests",
                 "splits_listed", "splits_visited_solutions", "splits_visited_early_exit", "splits_visited_no_split",
                 "filter_rounds", "sieved_rounds"]


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
    sum_limit, M = WORKLOADS[workload][:2]
    dmap = ffxpy.sieve_map(M)
    arr = np.zeros(16, dtype=np.uint64)
    ffxpy.check(lib.ffx_probe_counters(arr.ctypes.data, 1))
    sol = ffxpy.search(dmap, M, 5, sum_limit)
    ffxpy.check(lib.ffx_probe_counters(arr.ctypes.data, 1))
    res = {n: int(arr[i]) for i, n in enumerate(COUNTER_NAMES)}
    res["map_reads"] = res["reads_2k_plus_do"] + res["reads_d_plus_o2k"] + res["reads_o"] + res["reads_in_allbutone"]
    res["solutions"] = int(len(sol))
    print(json.dumps({workload: res}))


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        m = re.search(r"Model name:\s*(.+)", out)
        if m:
            return m.group(1).strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads():
    """Threads the host lets this process use (the GPU box's share, not the
    machine's core count)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def timed_run(cmd, timeout_s):
    with tempfile.TemporaryDirectory() as d:
        t0 = time.time()
        p = subprocess.run(cmd, cwd=d, capture_output=True, timeout=timeout_s)
        wall = time.time() - t0
    out = p.stdout.decode(errors="replace")
    pt = int(re.search(r"Prime time: (\d+)", out).group(1))
    ft = int(re.search(r"Freudenthal time: (\d+)", out).group(1))
    return pt * 1e-6, ft * 1e-6, wall


def cpu_baseline(workload, timeout_s, sample_k):
    """The oracle (oracle/ffref, the ff_seg algorithm restated in C) on a
    bounded, stratified sample of the same workload: the whole sieve of the
    workload's map (as ff_seg does it) and every sample_k-th odd sum of the
    search; the search time is scaled by sample_k to the whole range."""
    exe = os.path.join(ROOT, "oracle", "ffref")
    if not os.path.exists(exe):
        return None
    sum_limit, M = WORKLOADS[workload][:2]
    threads = min(16, host_threads())
    try:
        pt, ft, wall = timed_run([exe, "--threads", str(threads), "--search-sample", str(sample_k), "5",
                                  str(sum_limit)], timeout_s)
    except subprocess.TimeoutExpired:
        return {"value": None, "unit": "sums/s", "cores": threads, "kind": "port",
                "sample": f"timed out after {timeout_s} s"}
    nodd = (sum_limit - 5) // 2 + 1
    search_est = ft * sample_k
    map_bytes = (M >> 4) + 1
    return {
        "value": nodd / (pt + search_est),
        "unit": "sums/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle/ffref ({threads} threads) on 5..{sum_limit}: the whole sieve ({map_bytes} map bytes, "
                  f"{pt:.2f} s) and every {sample_k}th odd sum of the search ({ft:.2f} s, x{sample_k} = "
                  f"{search_est:.1f} s for all {nodd}); wall {wall:.1f} s; host {cpu_model()}",
        "sieve_s": round(pt, 3),
        "sieve_gbs": round(map_bytes / pt / 1e9, 3),
        "search_s_estimated": round(search_est, 2),
        "search_sums_per_s": nodd / search_est,
        "host_threads": host_threads(),
        "cpu_model": cpu_model(),
    }


def cpu_full(workload, reps, timeout_s):
    """The reference program itself (oracle/_ref/ff_seg built from the
    reference's source) on the whole workload, as the reference's own bench
    times it (scripts/bench_per_device.sh:56-108): its 31 threads and the
    host's thread count, median of `reps` runs, its own Prime/Freudenthal
    clocks."""
    sum_limit, M = WORKLOADS[workload][:2]
    nodd = (sum_limit - 5) // 2 + 1
    res = {"workload": WORKLOADS[workload][3], "odd_sums": nodd, "cpu_model": cpu_model(),
           "host_threads": host_threads(), "runs": {}}
    for name, threads in (("ff_seg", 31), ("ff_seg_t16", 16)):
        exe = os.path.join(ROOT, "oracle", "_ref", name)
        if not os.path.exists(exe):
            continue
        runs = []
        for r in range(reps):
            pt, ft, wall = timed_run([exe, "5", str(sum_limit)], timeout_s)
            runs.append({"prime_s": pt, "freudenthal_s": ft, "wall_s": round(wall, 2)})
            print(f"[cpu-full] {name} run {r + 1}: prime {pt:.2f} s, search {ft:.2f} s", file=sys.stderr, flush=True)
        tot = sorted(x["prime_s"] + x["freudenthal_s"] for x in runs)
        med = tot[len(tot) // 2]
        res["runs"][name] = {"threads": threads, "runs": runs, "median_s": round(med, 3),
                             "sums_per_s": nodd / med,
                             "median_prime_s": sorted(x["prime_s"] for x in runs)[len(runs) // 2],
                             "median_freudenthal_s": sorted(x["freudenthal_s"] for x in runs)[len(runs) // 2]}
    print(json.dumps(res))


def mode_leg(args, timeout_s):
    """One batched or served run of ff_sieve (its own process, its own
    device buffers, its output checked against plain slots); returns its
    JSON summary line."""
    exe = os.path.join(ROOT, "ff-gpu_amd", "bin", "ff_sieve")
    print(f"[bench] leg: ff_sieve {' '.join(args)}", file=sys.stderr, flush=True)
    try:
        p = subprocess.run([exe] + args, capture_output=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s} s", "parity": "MISMATCH"}
    try:
        res = json.loads(p.stdout.decode().strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": p.stderr.decode(errors="replace")[-500:], "parity": "MISMATCH"}
    if p.returncode != 0:
        res["error"] = p.stderr.decode(errors="replace")[-500:]
        res["parity"] = "MISMATCH"
    return res


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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sums2p21", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-timeout", type=float, default=240.0)
    ap.add_argument("--cpu-sample", type=int, default=32, help="the CPU leg searches every K-th odd sum")
    ap.add_argument("--no-modes", action="store_true", help="skip the batched and served legs")
    ap.add_argument("--count-probes", action="store_true",
                    help="only count the search's map reads by kind (needs make -C ff-gpu_amd counters)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only run the reference's full CPU runs (ff_seg, 31 and 16 threads, median of 3)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only copy a BYTES device buffer three times: the run that scales the HBM byte counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.count_probes:
        count_probes(args.workload)
        return
    if args.cpu_full:
        cpu_full(args.workload, 3, 900)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the sieve and search are single-device here (each GPU would "
              "re-sieve locally and search a disjoint sum range); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ffxpy
    from conftest import golden_sha  # noqa: E402  (the golden hash table only)

    sum_limit, M, leg, desc = WORKLOADS[args.workload]
    lib = ffxpy.load()
    ffxpy.init(0)
    dev = ffxpy.device_info(0)
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

    print(f"[bench] {args.workload}: {args.warmup} warmup + {args.steps} timed steps", file=sys.stderr, flush=True)
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
    for b in (dmap, slots, out, cnt):
        b.close()
    sp.close()
    qp.close()

    # The search's roofline.  Its HBM work is the map reads it makes (the
    # counters build counts them per launch), each a random single byte that
    # moves one whole 128-byte line; the reference-order probes it skips are
    # reported as work avoided.  Beside HBM: the measured ceiling of random
    # byte reads over a buffer of the map's size, and VALU issue (pmc run).
    pc = profile_json("search_probes.json").get(args.workload)
    ref_work = profile_json("reference_work.json").get(args.workload, {})
    traffic = profile_json("traffic.json").get(args.workload, {})
    issue = profile_json("issue.json").get(args.workload, {})
    ceiling = random_read_ceiling(map_bytes)
    roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
            "traffic": traffic.get("search_sums"), "kernel": "search_sums"}
    if pc:
        reads = pc["map_reads"]
        ach = reads * LINE_BYTES / (search_avg * 1e-3) / 1e9
        roof.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                     "algorithmic": f"{reads} map reads per launch (counted, profiles/search_probes.json) x "
                                    f"{LINE_BYTES} B, the line one random single-byte read moves"})
        if ref_work.get("probes"):
            roof["work_avoided"] = {"reference_order_probes": ref_work["probes"], "map_reads": reads,
                                    "ratio": round(ref_work["probes"] / reads, 2)}
        if ceiling:
            rate = reads / (search_avg * 1e-3)
            roof["random_read_roof"] = {"achieved_reads_per_s": round(rate, -6),
                                        "ceiling_reads_per_s": round(ceiling["reads_per_s"], -6),
                                        "frac": round(rate / ceiling["reads_per_s"], 3),
                                        "ceiling_from": f"bin/ff_random_reads over {map_bytes} bytes "
                                                        "(the map's size), best of 3"}
    if traffic.get("search_sums"):
        roof["traffic_gbs"] = round(traffic["search_sums"] / (search_avg * 1e-3) / 1e9, 1)
        roof["note"] = ("reads of small values recur across sums and are partly served from cache (measured traffic "
                        "vs counted reads x 128 B in profiles/traffic.json), so the rate against the ceiling of "
                        "uniform random reads can exceed 1")
    if issue.get("search_sums"):
        roof["valu_issue_frac"] = issue["search_sums"]["valu_per_cu_clock"]
    sieve_gbs = map_bytes / (sieve_avg * 1e-3) / 1e9
    roof_sieve = {"bound": "hbm", "achieved": round(sieve_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(sieve_gbs / HBM_PEAK_GBS, 4), "traffic": traffic.get("sieve_segments"),
                  "kernel": "sieve_segments", "algorithmic": f"{map_bytes} map bytes written once"}
    if issue.get("sieve_segments"):
        roof_sieve["valu_issue_frac"] = issue["sieve_segments"]["valu_per_cu_clock"]
    lds = profile_json("lds_floor.json").get(args.workload)
    if lds:
        roof_sieve["lds_strike_floor_ms"] = lds["floor_ms"]
        roof_sieve["frac_of_lds_floor"] = round(lds["floor_ms"] / sieve_avg, 3)
        if "floor_scattered_ms" in lds:
            # the strikes at the rate of scattered addresses, the rate the
            # strike patterns measure at (profiles/lds_atomics.md)
            roof_sieve["lds_strike_floor_scattered_ms"] = lds["floor_scattered_ms"]
            roof_sieve["frac_of_lds_floor_scattered"] = round(lds["floor_scattered_ms"] / sieve_avg, 3)
    roof_sieve["binding"] = ("LDS atomic ORs of the strikes, not HBM: the map is written once (traffic / algorithmic "
                             "bytes 1.00); frac_of_lds_floor* are the strikes' time at the measured ds_or_b32 rates "
                             "(profiles/lds_atomics.md) over the kernel's")
    roof["binding"] = "random single-byte map reads (random_read_roof) and VALU issue"

    cpu = None if args.no_cpu else cpu_baseline(args.workload, args.cpu_timeout, args.cpu_sample)
    cpu_ref = profile_json("cpu_baseline.json").get(args.workload)
    if cpu and cpu_ref:
        cpu["reference_full_runs"] = {k: {"threads": v["threads"], "median_s": v["median_s"],
                                          "sums_per_s": round(v["sums_per_s"], 1)}
                                      for k, v in cpu_ref.get("runs", {}).items()}
        cpu["reference_full_runs"]["_from"] = "profiles/cpu_baseline.json (recorded by bench.py --cpu-full)"
    legs = {}
    if not args.no_modes:
        legs["batched"] = mode_leg(["--batch", "--opt", "reps=3", "--opt", "verify=1"], 400)
        legs["served"] = mode_leg(["--serve", "--opt", "warm=1", "--opt", "plain_sample=100", "--opt", "verify=1"],
                                  400)
        legs["served_on_demand"] = mode_leg(["--serve", "--opt", "verify=1"], 400)
        legs["served_capacity"] = mode_leg(["--serve", "--opt", "warm=1", "--opt", "rate=12000", "--opt",
                                            "verify=1"], 400)
    legs_ok = all(l.get("parity") not in ("MISMATCH", None) and "error" not in l for l in legs.values())
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
        # `roofline` is the dominant kernel's (the longer average launch in
        # this run), the other kernel's beside it
        "roofline": roof_sieve if sieve_avg >= search_avg else roof,
        "roofline_other": roof if sieve_avg >= search_avg else roof_sieve,
        "sieve_ms": round(sieve_avg, 4),
        "sieve_gbs": round(sieve_gbs, 1),
        "search_ms": round(search_avg, 4),
        "search_sums_per_s": nodd / (search_avg * 1e-3),
        "compact_ms": round(sum(compact_ms) / len(compact_ms), 4),
        "solutions": c,
        "parity": "sha256 of emitted stdout == reference golden" if parity else "MISMATCH",
        "cpu_baseline": cpu,
    }
    res.update(legs)
    if cpu and cpu.get("value"):
        res["gpu_over_cpu"] = round(res["value"] / cpu["value"], 1)
    print(json.dumps(res))
    if not parity or not legs_ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
