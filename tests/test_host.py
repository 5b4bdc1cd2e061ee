"""Host logic of ff_sieve without a GPU: geometry, headers, budget, config,
ff_seg's chunk split, line format, workloads (ff-gpu_amd/bin/ff_hostcheck).
Known-answer values follow test/source/budget_selftest.cpp:43-181."""
import json
import os
import random
import subprocess

import pytest

from conftest import GOLDEN, ROOT, golden_headers, golden_solutions, run_ffref, tags_to_flags

HC = os.path.join(ROOT, "ff-gpu_amd", "bin", "ff_hostcheck")


@pytest.fixture(autouse=True, scope="module", params=["plain", "sanitized"])
def hostcheck_binary(request):
    """Every check runs twice: with the plain build and with the build under
    AddressSanitizer and UBSan, which aborts on any finding."""
    global HC
    name = "ff_hostcheck" if request.param == "plain" else "ff_hostcheck_san"
    HC = os.path.join(ROOT, "ff-gpu_amd", "bin", name)
    if not os.path.exists(HC):
        pytest.skip(f"{name} not built (make -C ff-gpu_amd)")
    yield HC


def hc(*args, env=None, check=True):
    p = subprocess.run([HC] + [str(a) for a in args], capture_output=True, text=True, env=env)
    if check:
        assert p.returncode == 0, (args, p.stderr)
    return p


def kv(*args):
    out = {}
    for line in hc(*args).stdout.splitlines():
        k, _, v = line.partition(" ")
        out.setdefault(k, []).append(v)
    return {k: (v[0] if len(v) == 1 else v) for k, v in out.items()}


def test_leg_geometry_kats():
    g = kv("leg", 5, 2097152)
    assert g["maxGeneratedPrime"] == "1048576"
    assert g["productLimit"] == "1099511627776"
    assert g["primeLimit"] == "274877906946"
    assert g["mapBytes"] == "17179869185"
    assert g["maxPrimeMapValue"] == "274877906959"
    assert kv("leg", 5, 1048576)["mapBytes"] == "4294967297"
    assert kv("leg", 5, 65536)["mapBytes"] == "16777217"
    assert kv("leg", 6, 100)["sumStart"] == "7"
    assert hc("workspace", 2097152).stdout.strip() == "24960"


def test_seg_headers_match_goldens():
    for leg, (h1, h2) in golden_headers().items():
        # sumLimit leg-1 has the same geometry as leg (same (L+1)>>1)
        for L in (leg, leg - 1):
            g = kv("seg", 5, L)
            assert g["header"] == [h1, h2], L


def test_seg_headers_match_ff_seg_cases():
    for c in json.load(open(os.path.join(GOLDEN, "seg_cases.json"))):
        args = list(c["args"]) + ([f"primemem={c['primemem']}"] if c["primemem"] else [])
        g = kv("seg", *args) if c["args"] else kv("seg", "primemem=" + str(c["primemem"] or 91268055040))
        want = [l for l in c["stdout"].split("\n") if l and ") sum" not in l and "time:" not in l
                and "input parameters" not in l]
        hdr = g["header"] if isinstance(g["header"], list) else [g["header"]]
        if c["rc"] & 255:
            assert g["fatal"] == "1" and hdr[0] == want[0]
        else:
            assert hdr == want, c["args"]


def test_capped_geometry_2_23():
    g = kv("seg", 5, 8388608)
    assert g["capped"] == "1"
    assert g["header"][1].startswith("numPrimesRequested ")
    assert g["pmapSumLimit"] == "4833682"


@pytest.mark.parametrize("rng_seed", [1, 2])
def test_chunks_match_oracle(ffref, rng_seed):
    # the host split equals the oracle's (which ff_seg's own stdout pins)
    rng = random.Random(rng_seed)
    cases = [(5, n) for n in range(5, 200)] + [(a, a + rng.randrange(0, 400)) for a in
                                               (rng.randrange(5, 10 ** 6) for _ in range(100))]
    for a, b in cases:
        want = run_ffref(["--chunks", a, b]).stdout.decode()
        assert hc("chunks", a | 1, b).stdout == want, (a, b)


def test_budget_kats():
    free = 34182397952
    b = kv("budget", free, 0.90)
    assert int(b["freeUsed"]) == int(0.90 * free)
    assert int(b["scratch"]) == 1 << 30
    assert int(b["backing"]) == (int(b["budget"]) - (1 << 30) - (64 << 20)) // 4096 * 4096
    b = kv("budget", free, 0.90, "cap=21474836480")
    assert int(b["budget"]) == 21474836480
    assert int(b["backing"]) == 21474836480 - (1 << 30) - (64 << 20)
    assert int(b["slabCount"]) == 18
    b = kv("budget", free, 0.90, "cap=8589934592")
    assert int(b["slabCount"]) == 6
    assert 2 * int(b["backing"]) < 17179869185  # two 8 GiB caps do not hold the 2^21 map
    # one shrink step toward the 0.50 floor
    s1 = kv("budget", 21474836480, 0.90, "shrink")
    assert abs(float(s1["fraction"]) - 0.70) < 1e-9
    s2 = kv("budget", 21474836480, float(s1["fraction"]), "shrink")
    assert abs(float(s2["fraction"]) - 0.60) < 1e-9
    assert int(s2["backing"]) < int(s1["backing"])
    fl = kv("budget", 21474836480, 0.50, "shrink")
    assert float(fl["fraction"]) == 0.50


def test_tier_plan():
    gib = 1 << 30
    m21 = 17179869185 * 16 - 1  # the 2^21 leg's map value: 16 GiB + 1 byte of map
    # the map fits the backing: no spill, the search keeps the whole map
    t = kv("tier", 20 * gib, 0, m21, gib)
    assert (t["gatePass"], t["spill"], int(t["searchMapValue"])) == ("1", "0", m21)
    # a 2 GiB backing fails alone (source/main.cpp:296-334) and passes with a host tier
    t = kv("tier", 2 * gib, 0, m21, 512 << 20)
    assert (t["gatePass"], t["spill"]) == ("0", "1")
    t = kv("tier", 2 * gib, 15 * gib, m21, 512 << 20)
    assert int(t["capacity"]) == 17 * gib and t["gatePass"] == "1" and t["spill"] == "1"
    # three 512 MiB slab buffers fit; the search's device map fills the backing
    assert (t["ring"], int(t["ringSlabBytes"])) == ("3", 512 << 20)
    assert int(t["searchMapValue"]) == 2 * gib * 16 - 1
    # one byte short of the map is still a failure
    t = kv("tier", 2 * gib, 17179869185 - 2 * gib - 1, m21, 512 << 20)
    assert t["gatePass"] == "0"
    # a slab larger than half the backing: two buffers of half, whole pages
    t = kv("tier", 15 << 20, 1 << 40, 2 ** 30 - 1, 1 << 30)
    assert (t["ring"], int(t["ringSlabBytes"])) == ("2", (15 << 19) // 4096 * 4096)
    # the capacity saturates instead of wrapping
    t = kv("tier", 2 * gib, 2 ** 64 - 1, m21, gib)
    assert int(t["capacity"]) == 2 ** 64 - 1 and t["gatePass"] == "1"


def test_bytes_to_human():
    # the units and decimals of the reference's messages (source/budget.cpp:53-68)
    for n, want in [(0, "0 B"), (1023, "1023 B"), (1024, "1.0 KiB"), (1536, "1.5 KiB"), ((1 << 20) - 1, "1024.0 KiB"),
                    (1 << 20, "1.0 MiB"), (17179869185, "16.00 GiB"), (3 << 29, "1.50 GiB"), (2 ** 64 - 1,
                                                                                           "17179869184.00 GiB")]:
        assert hc("human", n).stdout.strip() == want, n


def test_size_and_fraction_parsers():
    for s, v in [("20GiB", 21474836480), ("8GiB", 8589934592), ("1GiB", 1 << 30), ("1024MiB", 1 << 30),
                 ("64MiB", 67108864), ("4096", 4096), ("1.5GiB", 1610612736), ("0", 0)]:
        assert hc("size", s).stdout.strip() == str(v)
    for s in ["auto", "", "abc", "-5GiB", "20GiBX"]:
        assert hc("size", s, check=False).returncode != 0
    assert float(hc("fraction", "0.90").stdout) == 0.9
    assert hc("fraction", "x", check=False).returncode != 0
    assert hc("devfrac", "amd=0.001").stdout.split() == ["amd", "0.001"]
    for bad in ["amd", "amd=0.001,", "=0.5"]:
        assert hc("devfrac", bad, check=False).returncode != 0
    assert hc("human", 17179869185).stdout.strip() == "16.00 GiB"
    assert hc("human", 1 << 30).stdout.strip() == "1.00 GiB"


def test_config_validation_and_env():
    env = {k: v for k, v in os.environ.items() if not k.startswith("FF_")}
    assert hc("config", env=env).returncode == 0
    assert hc("config", "--vram-fraction", "2", env=env, check=False).returncode != 0
    assert hc("config", "--device-vram-fraction", "amd=0.001", env=env, check=False).returncode != 0
    assert hc("config", "--vram-budget", "0", env=env, check=False).returncode != 0
    assert hc("config", "--slab-size", "3", env=env, check=False).returncode != 0
    assert hc("config", "--bogus", env=env, check=False).returncode != 0
    assert hc("config", "--no-gpu", env=env, check=False).returncode != 0
    # CLI overrides env; malformed env is a hard error
    e2 = dict(env, FF_VRAM_BUDGET="20GiB")
    out = hc("config", env=e2).stdout
    assert "budgetCap 21474836480" in out
    out = hc("config", "--vram-budget", "8GiB", env=e2).stdout
    assert "budgetCap 8589934592" in out
    assert hc("config", env=dict(env, FF_VRAM_BUDGET="lots"), check=False).returncode != 0
    assert hc("config", env=dict(env, FF_DISABLE_DEVICE="intel"), check=False).returncode != 0
    assert hc("config", env=dict(env, FF_PULL_WEIGHTS="amd=0"), check=False).returncode != 0
    assert "primemem 3000" in hc("config", "--primemem", "3000", env=env).stdout


def test_line_format_matches_goldens():
    # every line of the 65536 golden, rebuilt by the C++ formatter
    rows = [r for r in golden_solutions() if r[0] <= 65536]
    from ffxpy import format_line
    for i, (s, lo, t) in enumerate(rows[:400], 1):
        line = hc("format", i, s, lo, tags_to_flags(t), 0).stdout
        assert line == format_line(i, s, lo, s - lo, tags_to_flags(t))


def test_workloads_are_seeded_and_shaped():
    a = hc("batch", 20261203, 4096).stdout.split("\n")
    assert a == hc("batch", 20261203, 4096).stdout.split("\n")
    rows = [tuple(map(int, l.split())) for l in a if l]
    assert len(rows) == 4096
    for s, l in rows:
        assert 5 <= s < (1 << 22) and (1 << 8) <= l - s + 1 <= (1 << 16)
    r = [l.split() for l in hc("served", 20261204, 2000).stdout.split("\n") if l]
    t = [float(x[2]) for x in r]
    assert t == sorted(t)
    for a_, b_, _ in r:
        a_, b_ = int(a_), int(b_)
        assert (1 << 18) <= b_ <= (1 << 23) and a_ >= 5 and b_ - a_ + 1 <= (1 << 18)


def _spans_bruteforce(spans, covered, top):
    a = set()
    for lo, hi in spans:
        a.update(range(lo, hi + 1, 2))
    c = set()
    for lo, hi in covered:
        c.update(range(lo, hi + 1, 2))
    return a, a - c


def _runs(sums):
    """Sorted odd sums -> maximal [lo, hi] runs of consecutive odd sums."""
    out = []
    for s in sorted(sums):
        if out and s == out[-1][1] + 2:
            out[-1][1] = s
        else:
            out.append([s, s])
    return [tuple(x) for x in out]


@pytest.mark.parametrize("rng_seed", [3, 4])
def test_interval_union_subtract_and_slots(rng_seed):
    # the batched/served modes share one slot per distinct odd sum: the
    # union, its slot offsets and the still-missing spans against sets
    rng = random.Random(rng_seed)
    for case in range(40):
        def span():
            lo = 2 * rng.randrange(2, 400) + 1
            return lo, lo + 2 * rng.randrange(0, 60 if case % 3 else 2)
        spans = [span() for _ in range(rng.randrange(1, 30))]
        covered = [span() for _ in range(rng.randrange(0, 20))]
        top = max(h for _, h in spans)
        feed = f"{len(spans)} {len(covered)}\n" + "".join(f"{a} {b}\n" for a, b in spans + covered)
        out = subprocess.run([HC, "spans"], input=feed, capture_output=True, text=True, check=True).stdout
        u_txt, m_txt, s_txt = out.split("--\n")
        allsums, missing = _spans_bruteforce(spans, covered, top)
        u = [tuple(map(int, l.split())) for l in u_txt.splitlines()]
        assert [(a, b) for a, b, _ in u] == _runs(allsums)
        off = 0
        for a, b, o in u:
            assert o == off
            off += (b - a) // 2 + 1
        assert [tuple(map(int, l.split())) for l in m_txt.splitlines()] == _runs(missing)
        order = sorted(allsums)
        slots = [int(x) for x in s_txt.split()]
        for i, s in enumerate(range(1, top + 7, 2)):
            assert slots[i] == (order.index(s) if s in allsums else -1), (case, s)


@pytest.mark.parametrize("a,b", [(4, 6), (5, 5), (1, 100), (100, 50), (7, 2627), (65535, 65535),
                                 (65536, 65600), (1000, 1003), (2, 9)])
def test_job_span_matches_oracle_chunks(ffref, a, b):
    # the odd sums a batched range evaluates are exactly those of ff_seg's
    # chunks (sumLimit+1 and skipped sums of tiny ranges included)
    j = kv("job", a, b)
    chunks = [tuple(map(int, c.split())) for c in (j["chunk"] if isinstance(j["chunk"], list) else [j["chunk"]])]
    ref = [tuple(map(int, l.split())) for l in run_ffref(["--chunks", a, b]).stdout.decode().splitlines()]
    assert chunks == ref
    odd = [s for lo, hi in ref for s in range(lo | 1, hi + 1, 2)]
    assert j["any"] == "1" and tuple(map(int, j["span"].split())) == (min(odd), max(odd))


def test_serve_admission_against_the_budget():
    # the serving loop's resident map: ff_seg's uncapped bound for the
    # largest sum when the backing holds it (2^23 - 1: ((2^22)^2)/4 + 2 rounded
    # up to a whole map byte), else the largest map the backing holds after
    # the fixed buffers, else refusal; --map-cap bounds it
    GiB, MiB = 1 << 30, 1 << 20
    assert hc("servemap", 512 * GiB, 16 * MiB, 2 ** 23 - 1, 0).stdout.strip() == "4398046511119"
    M = int(hc("servemap", 256 * GiB, 16 * MiB, 2 ** 23 - 1, 0).stdout)
    room = 256 * GiB - 16 * MiB
    assert room - 16 < (M >> 4) + 1 <= room and M % 16 == 15
    assert hc("servemap", 64 * MiB, 16 * MiB, 2 ** 23 - 1, 0).stdout.strip() == "0"
    assert hc("servemap", 512 * GiB, 16 * MiB, 2 ** 23 - 1, 2 ** 34 - 1).stdout.strip() == str(2 ** 34 - 1)
    # small sums: the uncapped bound of 1048575 is ff_seg's own M for that leg
    assert hc("servemap", 512 * GiB, 16 * MiB, 1048575, 0).stdout.strip() == "68719476751"


def test_sha256_matches_hashlib():
    # the digest the batched and served modes write per range
    import hashlib
    rng = random.Random(41)
    for n in [0, 1, 3, 55, 56, 63, 64, 65, 119, 120, 4095, 100003]:
        data = bytes(rng.getrandbits(8) for _ in range(n))
        p = subprocess.run([HC, "sha256"], input=data, capture_output=True)
        assert p.returncode == 0
        assert p.stdout.decode().strip() == hashlib.sha256(data).hexdigest(), n


@pytest.mark.parametrize("a,b", [(1001, 3000), (4, 6), (5, 5), (100, 50), (60001, 70000), (5, 2627),
                                 (131001, 131071), (262001, 262100)])
def test_range_digest_matches_oracle_stdout(ffref, a, b):
    # the emitter and digest of the batched and served modes, fed the
    # oracle's per-sum answers, hash to the oracle's own ff_seg stdout
    import hashlib
    from conftest import normalize
    want = hashlib.sha256(normalize(run_ffref([a, b]).stdout.decode()).encode()).hexdigest()
    first, last = max(5, a | 1), max(b, max(5, a | 1))
    sums = "\n".join(str(s) for s in range(first, last + 2, 2)).encode()
    solved = run_ffref(["--solve"], stdin=sums).stdout
    p = subprocess.run([HC, "rangedigest", str(a), str(b)], input=solved, capture_output=True)
    assert p.returncode == 0, p.stderr
    got = dict(l.split(" ", 1) for l in p.stdout.decode().splitlines())
    assert got["digest"] == want, (a, b)
