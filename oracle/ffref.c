/*
 * ffref -- CPU oracle for the ff_seg prime-bitmap sieve and Freudenthal search.
 *
 * TEST INFRASTRUCTURE ONLY.  This program is the checker the parity tests,
 * the smoke entry and the bench's CPU-baseline leg compare against.  No
 * product code links, loads or calls it; the product path (libffx + ff_sieve)
 * fails loudly instead of falling back to it.
 *
 * It restates the behaviour of reference/segmentedSieve.C ("ff_seg") in plain
 * C with pthreads.  Each function cites the reference lines it follows:
 *
 *   geometry + header lines        segmentedSieve.C:846-855, :874-879, :136-182
 *   odd-only bitmap sieve          segmentedSieve.C:177-222, :255-270
 *   IsPrime / Miller-Rabin          segmentedSieve.C:283-296, :352-387
 *   PTP / AllButOne... / Sammy      segmentedSieve.C:419-477, :529-639
 *   per-sum filter + line format    segmentedSieve.C:800-823, :645-677
 *   thread partition + merge        segmentedSieve.C:706-798, :905-912
 *
 * How each part is pinned (tests/test_oracle.py):
 *   - full stdout sha256 == goldens/SHA256SUMS for legs 65536..2097152
 *     (the 2^21 leg by hash only) and per-solution fixtures in tests/golden;
 *   - bitmap: popcount == pi(M) - 1 (known pi values), brute force on small
 *     ranges, byte-compare with the reference binary's --dump-map when it is
 *     built (oracle/_ref);
 *   - Miller-Rabin: against trial division and against the bitmap;
 *   - tiny ranges and the capped (PRIMEMEM) geometry: against oracle/_ref.
 *
 * Modes
 *   ffref [opts] [sumStart] sumLimit     ff_seg-identical stdout
 *   ffref --solve [opts] < sums          "sum low high" per line for each sum
 *                                        read from stdin (pure F(s); "sum 0 0"
 *                                        when the sum is not emitted)
 * Options
 *   --threads N        worker threads (default: online CPUs)
 *   --primemem BYTES   the reference's PRIMEMEM cap (default 85 GiB)
 *   --map-cap VALUE    (--solve only) largest value held in the bitmap;
 *                      Miller-Rabin answers above it
 *   --dump-map FILE    write the bitmap like ff_seg --dump-map and exit
 *   --counters         print reference-order work counters to stderr
 *   --search-sample K  search only every K-th odd sum (timing sample; the
 *                      output is then NOT ff_seg's)
 *   --chunks           print ff_seg's chunk list ("lo hi" per line, print
 *                      order) for the positional range and exit
 *   --mr               read odd n > 17 from stdin, print "n 0|1" from the
 *                      Miller-Rabin routine alone
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define SEG_PARTITION_THREADS 31u            /* segmentedSieve.C:11-13 */
#define SEG_PRIMEMEM_DEFAULT (85ull << 30)   /* segmentedSieve.C:8-9 */

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Exact floor(sqrt(n)); same contract as isqrt64, segmentedSieve.C:37-52. */
static uint64_t isqrt_u64(uint64_t n)
{
    uint64_t r = (uint64_t)sqrt((double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) ++r;
    return r;
}

/* double -> uint32 the way g++ on x86-64 converts in ff_seg (truncate via
 * a 64-bit signed conversion, keep the low 32 bits). */
static uint32_t dbl_to_u32(double x) { return (uint32_t)(int64_t)x; }

/* ------------------------------------------------------------------ */
/* geometry (segmentedSieve.C:846-855, :874-879, :136-182)              */
/* ------------------------------------------------------------------ */

typedef struct {
    uint32_t sumStart, sumLimit;
    uint64_t productLimit, primeLimit;
    uint64_t R, M, mapBytes, totalBytes, sqrtLimit;
    char header[4][160];
    int nheader;
    int fatal;            /* ff_seg would print the "more memory" message and exit(-1) */
} geom_t;

static void geom_compute(geom_t *g, uint32_t sumStart, uint32_t sumLimit, uint64_t primemem)
{
    memset(g, 0, sizeof *g);
    /* main :854-855 */
    if (sumStart < 5) sumStart = 5;
    if (sumStart > sumLimit) sumLimit = sumStart;
    g->sumStart = sumStart;
    g->sumLimit = sumLimit;
    /* main :875-879 */
    uint32_t mgp = (uint32_t)(sumLimit + 1u) >> 1;
    g->productLimit = (uint64_t)mgp * mgp;
    g->primeLimit = (g->productLimit >> 2) + 2;
    double xlogx = (double)mgp / log((double)mgp);
    uint32_t nPrimes = dbl_to_u32(xlogx + 1.2762 * xlogx / log((double)mgp));
    /* Prime::Prime :137-147 */
    uint64_t R = nPrimes > 1 ? nPrimes : 1;
    uint64_t limit = g->primeLimit;
    uint64_t mapBytes = ((limit > 1 ? limit : 2) + 15) >> 4;
    uint64_t M = (mapBytes << 4) - 1;
    uint64_t estMax = (uint64_t)((double)R * (log((double)R) + log(log((double)R))));
    if (M < estMax) M = estMax;
    mapBytes = (M + 15) >> 4;
    uint64_t reqBytes = 4 * R;
    uint64_t total = reqBytes + mapBytes;
    snprintf(g->header[g->nheader++], 160, "maxPrimeMapValue %llu numPrimesRequested %llu totalBytes %llu",
             (unsigned long long)M, (unsigned long long)R, (unsigned long long)total);
    if (total > primemem) {
        /* capped geometry :148-176 */
        if (16 * primemem > 3 * estMax) {
            uint64_t R2 = 2 * R / 5, req2 = 4 * R2;
            estMax = (uint64_t)((double)R2 * (log((double)R2) + log(log((double)R2))));
            if (req2 + (estMax + 1) / 2 > primemem) {
                R = ((reqBytes * primemem + total - 1) / total + 3) & ~(uint64_t)3;
                reqBytes = 4 * R;
                snprintf(g->header[g->nheader++], 160,
                         "You probably should consider how badly this run is going to perform, just sayin'.");
            } else {
                R = R2;
                reqBytes = req2;
            }
            mapBytes = primemem - reqBytes;
            total = reqBytes + mapBytes;
            M = (mapBytes << 4) - 1;
            snprintf(g->header[g->nheader++], 160, "numPrimesRequested %llu maxPrimeMapValue %llu totalBytes %llu",
                     (unsigned long long)R, (unsigned long long)M, (unsigned long long)total);
        } else {
            g->fatal = 1;
            return;
        }
    }
    g->R = R;
    g->M = M;
    g->mapBytes = mapBytes;
    g->totalBytes = total;
    g->sqrtLimit = isqrt_u64(M);
    /* the sqrtLimit line is printed once R and M are final (:182) */
    snprintf(g->header[g->nheader++], 160, "maxPrimeMapValue %llu numPrimesRequested %llu sqrtLimit %llu",
                 (unsigned long long)M, (unsigned long long)R, (unsigned long long)g->sqrtLimit);
}

/* ------------------------------------------------------------------ */
/* bitmap sieve (segmentedSieve.C:177-222, :255-270)                    */
/* Bit for odd n lives in byte n>>4 under mask 0x80>>((n>>1)&7).        */
/* ------------------------------------------------------------------ */

typedef struct {
    uint8_t *map;
    uint64_t mapBytes;
    uint64_t M;             /* largest value answered by the bitmap */
    uint32_t *tdPrimes;     /* odd primes 3,5,7,... for trial division */
    uint32_t nTd;
} ptab_t;

/* Odd primes <= lim, by a plain sieve (the "small map" of :184-203). */
static uint32_t *small_odd_primes(uint64_t lim, uint32_t *count)
{
    uint64_t n = lim + 1;
    uint8_t *comp = calloc(n + 1, 1);
    uint32_t cnt = 0;
    for (uint64_t p = 3; p * p <= lim; p += 2)
        if (!comp[p])
            for (uint64_t i = p * p; i <= lim; i += 2 * p) comp[i] = 1;
    for (uint64_t p = 3; p <= lim; p += 2)
        if (!comp[p]) ++cnt;
    uint32_t *out = malloc(sizeof(uint32_t) * (cnt ? cnt : 1));
    cnt = 0;
    for (uint64_t p = 3; p <= lim; p += 2)
        if (!comp[p]) out[cnt++] = (uint32_t)p;
    free(comp);
    *count = cnt;
    return out;
}

typedef struct {
    uint8_t *map;
    uint64_t lo, hi;          /* value range [lo, hi), lo a multiple of 16 */
    const uint32_t *primes;
    uint32_t nprimes;
    _Atomic uint64_t *next;   /* work counter over 2^20-value blocks */
    uint64_t vmax;            /* sieve values [0, vmax) */
} sieve_job_t;

/* Clear every odd multiple m >= p*p of each sieving prime in [lo, hi):
 * the strike loop of SegmentFill (:255-270), run per 1 Mi-value block. */
static void *sieve_worker(void *arg)
{
    sieve_job_t *j = arg;
    const uint64_t blk = 1ull << 20;
    for (;;) {
        uint64_t b = atomic_fetch_add(j->next, 1);
        uint64_t lo = b * blk;
        if (lo >= j->vmax) break;
        uint64_t hi = lo + blk < j->vmax ? lo + blk : j->vmax;
        for (uint32_t i = 0; i < j->nprimes; ++i) {
            uint64_t p = j->primes[i];
            if (p * p >= hi) break;
            uint64_t m = (lo + p - 1) / p * p;
            if (m < p * p) m = p * p;
            if (!(m & 1)) m += p;
            for (; m < hi; m += 2 * p) j->map[m >> 4] &= (uint8_t)~(0x80u >> ((m >> 1) & 7));
        }
    }
    return NULL;
}

/* Builds the bitmap for values [0, 16*mapBytes): all ones, value 1 cleared,
 * odd composites cleared (Prime::Prime :177-219). */
static void sieve_build(ptab_t *pt, uint64_t mapBytes, uint64_t M, int threads)
{
    pt->mapBytes = mapBytes;
    pt->M = M;
    pt->map = malloc(mapBytes ? mapBytes : 1);
    if (!pt->map) {
        fprintf(stderr, "ffref: cannot allocate %llu-byte bitmap\n", (unsigned long long)mapBytes);
        exit(3);
    }
    memset(pt->map, 0xFF, mapBytes);
    pt->map[0] ^= 0x80;
    uint64_t vmax = mapBytes << 4;
    uint64_t sq = isqrt_u64(M);
    uint32_t np;
    uint32_t *primes = small_odd_primes(sq, &np);
    _Atomic uint64_t next = 0;
    sieve_job_t job = {pt->map, 0, 0, primes, np, &next, vmax};
    pthread_t th[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, sieve_worker, &job);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(primes);
}

/* ------------------------------------------------------------------ */
/* IsPrime (segmentedSieve.C:283-296) and Miller-Rabin (:352-387)       */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t splits, probes, mr_calls, ab1_calls, trial_divs;
} ctr_t;

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t n)
{
    return (uint64_t)((unsigned __int128)a * b % n);
}

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t n)
{
    uint64_t r = 1;
    b %= n;
    while (e) {
        if (e & 1) r = mulmod(r, b, n);
        b = mulmod(b, b, n);
        e >>= 1;
    }
    return r;
}

/* Strong-probable-prime test with the reference's base list {2..17} and the
 * base count chosen by magnitude (AskMillerRabin :364-387). */
static int miller_rabin(uint64_t n)
{
    static const uint64_t bases[7] = {2, 3, 5, 7, 11, 13, 17};
    unsigned r = (unsigned)__builtin_ctzll(n - 1);
    uint64_t d = (n - 1) >> r;
    int nb = n < 3215031751ull ? 4 : n < 2152302898747ull ? 5 : n < 3474749660383ull ? 6 : 7;
    for (int i = 0; i < nb; ++i) {
        uint64_t x = powmod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        unsigned j = 1;
        for (; j < r; ++j) {
            x = mulmod(x, x, n);
            if (x == 1) return 0;
            if (x == n - 1) break;
        }
        if (x != n - 1) return 0;
    }
    return 1;
}

static inline int is_prime(const ptab_t *pt, uint64_t n, ctr_t *c)
{
    if (n == 2) return 1;
    if (!(n & 1)) return 0;
    if (n > pt->M) {
        c->mr_calls++;
        return miller_rabin(n);
    }
    c->probes++;
    return (pt->map[n >> 4] >> (7 - ((n >> 1) & 7))) & 1;
}

/* PTP: ProductOfTermPairsHasSingleFactorPair (:419-422). */
static inline int ptp(const ptab_t *pt, uint64_t i, ctr_t *c)
{
    if (!(i & 1)) return 1;
    return is_prime(pt, i - 2, c);
}

/* ------------------------------------------------------------------ */
/* AllButOne... (:424-477)                                              */
/* Counts divisors f of P = q*2^k with f <= isqrt(P), f in              */
/* {2^k} u {d | q, d > 1} u {2^k*d : d | q, d > 1}, for which           */
/* PTP(f + P/f) is false; true iff exactly one.  Divisors 2^j*d with    */
/* 0 < j < k pair two even terms, PTP of an even sum is true, so the    */
/* reference never lists them; neither does this.                        */
/* ------------------------------------------------------------------ */

static int ab1(const ptab_t *pt, uint32_t k, uint64_t q, ctr_t *c)
{
    c->ab1_calls++;
    const uint64_t P = q << k, L = isqrt_u64(P), pk = 1ull << k;
    int bad = 0;
    if (pk <= L) bad += !ptp(pt, pk + q, c);
    /* factor q by trial division over odd primes */
    uint64_t pr[48];
    unsigned ex[48], nf = 0;
    uint64_t rem = q;
    for (uint32_t i = 0; i < pt->nTd; ++i) {
        uint64_t p = pt->tdPrimes[i];
        if (p * p > rem) break;
        c->trial_divs++;
        if (rem % p == 0) {
            unsigned e = 0;
            do { rem /= p; ++e; } while (rem % p == 0);
            pr[nf] = p; ex[nf] = e; ++nf;
        }
    }
    if (rem > 1) { pr[nf] = rem; ex[nf] = 1; ++nf; }
    /* odometer over exponent vectors: every divisor d > 1 of q once */
    unsigned e[48] = {0};
    uint64_t d = 1;
    for (;;) {
        unsigned i = 0;
        while (i < nf && e[i] == ex[i]) {
            for (unsigned t = 0; t < ex[i]; ++t) d /= pr[i];
            e[i] = 0;
            ++i;
        }
        if (i == nf) break;
        e[i]++;
        d *= pr[i];
        if (d <= L && !ptp(pt, d + P / d, c))
            if (++bad > 1) return 0;
        if (d <= L / pk && !ptp(pt, d * pk + P / (d * pk), c))
            if (++bad > 1) return 0;
    }
    return bad == 1;
}

/* DoesPeterKnow (:632-639). */
static int dpk(const ptab_t *pt, uint32_t k, uint64_t d, uint64_t o, uint64_t q, ctr_t *c)
{
    if (!ptp(pt, (1ull << k) + q, c)) return 0;
    if (d != o && !ptp(pt, d + (o << k), c)) return 0;
    return ab1(pt, k, q, c);
}

/* DoesSammyKnow (:529-639): exactly one valid split e + o = s among the
 * four families; the unique split is returned as (even, odd) terms. */
static int sammy_knows(const ptab_t *pt, uint32_t s, uint32_t *te, uint32_t *to, ctr_t *c)
{
    int cnt = 0;
    uint32_t ce = 0, co = 0;
#define FOUND(E, O) do { if (cnt == 1 && ce == 0) { ce = (E); co = (O); } } while (0)
    /* Power2Prime (:547-564): s = 2^k + o, o prime; composite o remembered */
    uint32_t compK[40];
    int ncomp = 0;
    for (uint32_t k = 2; (1u << k) < s - 2; ++k) {
        uint32_t o = s - (1u << k);
        c->splits++;
        if (is_prime(pt, o, c)) {
            if (++cnt > 1) break;
            FOUND(1u << k, o);
        } else {
            compK[ncomp++] = k;
        }
    }
    /* Power2Odd (:566-589): odd k from the largest with 5*2^k < s-2 down to 3 */
    if (cnt <= 1) {
        uint32_t kf = 2;
        while ((5u << kf) < s - 2) ++kf;
        for (uint32_t k = (kf - 2) | 1; k >= 2 && k < 64; k -= 2) {
            for (uint32_t d = 5; (d << k) < s - 2; d += 2) {
                uint32_t o = s - (d << k);
                uint64_t q = (uint64_t)d * o;
                c->splits++;
                if (q % 3 == 0) continue;
                cnt += dpk(pt, k, d, o, q, c);
                if (cnt > 1) break;
                FOUND(d << k, o);
            }
            if (cnt > 1) break;
        }
    }
    /* Power2Composite (:591-605): remembered k, largest first */
    if (cnt <= 1) {
        for (int i = ncomp - 1; i >= 0; --i) {
            uint32_t k = compK[i];
            uint32_t o = s - (1u << k);
            cnt += ab1(pt, k, o, c);
            if (cnt > 1) break;
            FOUND(1u << k, o);
        }
    }
    /* Power2Even (:607-630): even k from the largest with 3*2^k < s-2 down to 2 */
    if (cnt <= 1) {
        uint32_t kf = 2;
        while ((3u << kf) < s - 2) ++kf;
        for (uint32_t k = (kf - 1) & ~1u; k >= 2 && k < 64; k -= 2) {
            for (uint32_t d = 3; (d << k) < s - 2; d += 2) {
                uint32_t o = s - (d << k);
                uint64_t q = (uint64_t)d * o;
                c->splits++;
                if (q % 3 != 0 && !is_prime(pt, o, c)) continue;
                cnt += dpk(pt, k, d, o, q, c);
                if (cnt > 1) break;
                FOUND(d << k, o);
            }
            if (cnt > 1) break;
        }
    }
#undef FOUND
    *te = ce;
    *to = co;
    return cnt == 1;
}

/* Per-sum emit rule (RunIt :805-815): statement 1, then DoesSammyKnow.
 * Returns the low term (0 when the sum is not emitted). */
static uint32_t solve_sum(const ptab_t *pt, uint32_t s, uint32_t *high, ctr_t *c)
{
    *high = 0;
    if (ptp(pt, s, c)) return 0;
    if (s % 3 == 0 && !is_prime(pt, s / 3, c)) return 0;
    uint32_t a, b;
    if (!sammy_knows(pt, s, &a, &b, c)) return 0;
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    *high = hi;
    return lo;
}

/* ------------------------------------------------------------------ */
/* thread partition (GenerateThreads :706-761) and merge (:763-791)     */
/* ------------------------------------------------------------------ */

typedef struct {
    uint32_t lo, hi;  /* sums [lo|1, hi] */
} chunk_t;

/* Emits the chunk list in ff_seg's output order: every front chunk in
 * launch order, then every live back chunk in reverse launch order. */
static int seg_partition(uint32_t sumStart, uint32_t sumLimit, chunk_t *out)
{
    const uint32_t T = SEG_PARTITION_THREADS;
    uint32_t numSum = sumLimit - sumStart + 1;
    double delta = (double)numSum / (2 * T);
    uint32_t prev1 = 0, prev2 = numSum;
    int moved1 = 1, moved2 = 1;
    chunk_t front[SEG_PARTITION_THREADS], back[SEG_PARTITION_THREADS];
    int hasBack[SEG_PARTITION_THREADS];
    int n = 0;
    for (uint32_t t = 0; t < T; ++t) {
        uint32_t a = dbl_to_u32(t * delta + .5);
        if (a > prev1) { prev1 = a; moved1 = 1; }
        uint32_t t2 = 2 * T - t - 1;
        uint32_t b = dbl_to_u32(t2 * delta + .5);
        if (b < prev2) { prev2 = b; moved2 = 1; }
        if (!(moved1 && moved2)) continue;
        moved1 = moved2 = 0;
        uint32_t aL = dbl_to_u32((t + 1) * delta - .5);
        if (aL < a) aL = a;
        front[n].lo = a + sumStart;
        front[n].hi = aL + sumStart;
        hasBack[n] = a < b;
        if (hasBack[n]) {
            uint32_t bL = dbl_to_u32((t2 + 1) * delta - .5);
            if (bL < b) bL = b;
            back[n].lo = b + sumStart;
            back[n].hi = bL + sumStart;
        }
        ++n;
    }
    int m = 0;
    for (int i = 0; i < n; ++i) out[m++] = front[i];
    for (int i = n - 1; i >= 0; --i)
        if (hasBack[i]) out[m++] = back[i];
    return m;
}

/* ------------------------------------------------------------------ */
/* parallel search over a contiguous odd-sum range                       */
/* ------------------------------------------------------------------ */

typedef struct {
    const ptab_t *pt;
    uint32_t s0;              /* first odd sum */
    uint64_t nodd;            /* slots */
    uint32_t stride;          /* search every stride-th slot (1 = all) */
    uint32_t *low, *high;     /* per-slot result */
    _Atomic uint64_t next;    /* blocks handed out from the top (costly end first) */
    uint64_t nblk;
    ctr_t *ctr;               /* per thread */
    int nthreads;
    _Atomic int tid;
} search_job_t;

static void *search_worker(void *arg)
{
    search_job_t *j = arg;
    int me = atomic_fetch_add(&j->tid, 1);
    ctr_t c = {0};
    const uint64_t B = 256;
    for (;;) {
        uint64_t b = atomic_fetch_add(&j->next, 1);
        if (b >= j->nblk) break;
        uint64_t blk = j->nblk - 1 - b;
        uint64_t i0 = blk * B, i1 = i0 + B < j->nodd ? i0 + B : j->nodd;
        for (uint64_t i = i0; i < i1; ++i) {
            if (j->stride > 1 && i % j->stride) continue;
            uint32_t s = j->s0 + 2 * (uint32_t)i, hi;
            j->low[i] = solve_sum(j->pt, s, &hi, &c);
            j->high[i] = hi;
        }
    }
    j->ctr[me] = c;
    return NULL;
}

static void search_range(const ptab_t *pt, uint32_t s0, uint64_t nodd, uint32_t stride, uint32_t *low,
                         uint32_t *high, int threads, ctr_t *total)
{
    search_job_t j;
    memset(&j, 0, sizeof j);
    j.pt = pt;
    j.s0 = s0;
    j.nodd = nodd;
    j.stride = stride;
    j.low = low;
    j.high = high;
    j.nblk = (nodd + 255) / 256;
    j.nthreads = threads;
    j.ctr = calloc((size_t)threads, sizeof(ctr_t));
    pthread_t th[256];
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, search_worker, &j);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    memset(total, 0, sizeof *total);
    for (int t = 0; t < threads; ++t) {
        total->splits += j.ctr[t].splits;
        total->probes += j.ctr[t].probes;
        total->mr_calls += j.ctr[t].mr_calls;
        total->ab1_calls += j.ctr[t].ab1_calls;
        total->trial_divs += j.ctr[t].trial_divs;
    }
    free(j.ctr);
}

/* ------------------------------------------------------------------ */
/* output (OutputTags :663-677, RunIt :816-819, merge :772/:786)         */
/* ------------------------------------------------------------------ */

static int fmt_line(char *buf, const ptab_t *pt, uint64_t ord, uint32_t s, uint32_t lo, uint32_t hi)
{
    ctr_t c = {0};
    char tl[24], th[16];
    if (is_prime(pt, lo, &c)) {
        strcpy(tl, " (prime), ");
    } else if ((lo & (lo - 1)) == 0) {
        unsigned k = (unsigned)__builtin_ctz(lo);
        snprintf(tl, sizeof tl, " (2^%u),%s", k, k < 10 ? "   " : "  ");
    } else {
        strcpy(tl, ",         ");
    }
    if (is_prime(pt, hi, &c)) {
        strcpy(th, " (prime)");
    } else if ((hi & (hi - 1)) == 0) {
        snprintf(th, sizeof th, " (2^%u)", (unsigned)__builtin_ctz(hi));
    } else {
        th[0] = 0;
    }
    return sprintf(buf, "%7llu) sum =%9u, product =%16llu,  low term =%9u%shigh term =%9u%s\n",
                   (unsigned long long)ord, s, (unsigned long long)lo * hi, lo, tl, hi, th);
}

/* ------------------------------------------------------------------ */
/* trial-division prime list                                             */
/* ------------------------------------------------------------------ */

static void td_primes(ptab_t *pt, uint64_t maxQ)
{
    uint64_t lim = isqrt_u64(maxQ) + 1;
    pt->tdPrimes = small_odd_primes(lim, &pt->nTd);
}

/* ------------------------------------------------------------------ */
/* modes                                                                 */
/* ------------------------------------------------------------------ */

static int run_seg(uint32_t sumStart, uint32_t sumLimit, uint64_t primemem, int threads, const char *dumpMap,
                   int counters, uint32_t sample)
{
    geom_t g;
    double t0 = now_s();
    geom_compute(&g, sumStart, sumLimit, primemem);
    for (int i = 0; i < g.nheader && (i < 1 || !g.fatal); ++i) printf("%s\n", g.header[i]);
    if (g.fatal) {
        printf("\nThe input parameters would indicate that more memory is required. Please adjust input "
               "parameters or Prime::maxPrimeAlloc.\n\n");
        return 255;
    }
    fflush(stdout);
    ptab_t pt;
    memset(&pt, 0, sizeof pt);
    sieve_build(&pt, g.mapBytes, g.M, threads);
    double t1 = now_s();
    if (dumpMap) {
        FILE *f = fopen(dumpMap, "wb");
        if (!f || fwrite(pt.map, 1, pt.mapBytes, f) != pt.mapBytes) {
            fprintf(stderr, "ffref: cannot write %s\n", dumpMap);
            return 4;
        }
        fclose(f);
        return 0;
    }
    /* GenerateThreads is called once, or twice around pmapSumLimit (:905-912) */
    uint64_t pmapSumLimit = isqrt_u64((g.M - 2) << 2) << 1;
    uint32_t calls[2][2];
    int ncalls = 0;
    if (g.sumLimit <= pmapSumLimit) {
        calls[ncalls][0] = g.sumStart; calls[ncalls][1] = g.sumLimit; ++ncalls;
    } else {
        calls[ncalls][0] = g.sumStart; calls[ncalls][1] = (uint32_t)pmapSumLimit; ++ncalls;
        calls[ncalls][0] = (uint32_t)(pmapSumLimit + 1) | 1; calls[ncalls][1] = g.sumLimit; ++ncalls;
    }
    chunk_t chunks[2][2 * SEG_PARTITION_THREADS];
    int nch[2];
    uint32_t smin = 0xFFFFFFFFu, smax = 0;
    for (int c = 0; c < ncalls; ++c) {
        nch[c] = seg_partition(calls[c][0], calls[c][1], chunks[c]);
        for (int i = 0; i < nch[c]; ++i) {
            uint32_t lo = chunks[c][i].lo | 1, hi = chunks[c][i].hi;
            if (lo > hi) continue;
            if (lo < smin) smin = lo;
            if (hi > smax) smax = hi;
        }
    }
    uint64_t nodd = smax >= smin ? (uint64_t)(smax - smin) / 2 + 1 : 0;
    uint32_t *low = calloc(nodd ? nodd : 1, 4), *high = calloc(nodd ? nodd : 1, 4);
    td_primes(&pt, (uint64_t)smax * smax / 4 + 4);
    ctr_t ctr;
    memset(&ctr, 0, sizeof ctr);
    if (nodd) search_range(&pt, smin, nodd, sample ? sample : 1, low, high, threads, &ctr);
    /* merge in ff_seg order with global renumbering */
    uint64_t ord = 0;
    char line[256];
    for (int c = 0; c < ncalls; ++c)
        for (int i = 0; i < nch[c]; ++i)
            for (uint64_t s = chunks[c][i].lo | 1; s <= chunks[c][i].hi; s += 2) {
                uint64_t idx = (s - smin) / 2;
                if (!low[idx]) continue;
                fmt_line(line, &pt, ++ord, (uint32_t)s, low[idx], high[idx]);
                fputs(line, stdout);
            }
    double t2 = now_s();
    printf("Prime time: %llu \xce\xbcs\nFreudenthal time: %llu \xce\xbcs\n", (unsigned long long)((t1 - t0) * 1e6),
           (unsigned long long)((t2 - t1) * 1e6));
    if (counters)
        fprintf(stderr,
                "ffref counters: sums=%llu splits=%llu probes=%llu mr=%llu ab1=%llu trialdiv=%llu "
                "sieve_s=%.6f search_s=%.6f threads=%d mapBytes=%llu\n",
                (unsigned long long)nodd, (unsigned long long)ctr.splits, (unsigned long long)ctr.probes,
                (unsigned long long)ctr.mr_calls, (unsigned long long)ctr.ab1_calls,
                (unsigned long long)ctr.trial_divs, t1 - t0, t2 - t1, threads, (unsigned long long)g.mapBytes);
    free(low);
    free(high);
    free(pt.map);
    free(pt.tdPrimes);
    return 0;
}

typedef struct {
    const ptab_t *pt;
    const uint32_t *s;
    uint32_t *lo, *hi;
    size_t n;
    _Atomic size_t *next;
} solve_job_t;

static void *solve_worker(void *a)
{
    solve_job_t *j = a;
    ctr_t c = {0};
    for (;;) {
        size_t i = atomic_fetch_add(j->next, 1);
        if (i >= j->n) break;
        if (j->s[i] < 5 || !(j->s[i] & 1)) continue;
        j->lo[i] = solve_sum(j->pt, j->s[i], &j->hi[i], &c);
    }
    return NULL;
}

/* --solve: pure F(s) for an arbitrary list of sums.  The bitmap covers values
 * up to mapCap (default: what the largest sum's ff_seg geometry would hold,
 * capped at 2^33); Miller-Rabin answers above it, as IsPrime does. */
static int run_solve(int threads, uint64_t mapCap)
{
    size_t cap = 1024, n = 0;
    uint32_t *sums = malloc(cap * 4);
    unsigned long long v;
    while (scanf("%llu", &v) == 1) {
        if (n == cap) { cap *= 2; sums = realloc(sums, cap * 4); }
        sums[n++] = (uint32_t)v;
    }
    uint32_t smax = 0;
    for (size_t i = 0; i < n; ++i) if (sums[i] > smax) smax = sums[i];
    uint64_t want = (uint64_t)smax * smax / 16 + 16;
    if (mapCap == 0) mapCap = want < (1ull << 33) ? want : (1ull << 33);
    uint64_t mapBytes = (mapCap + 16) >> 4;
    ptab_t pt;
    memset(&pt, 0, sizeof pt);
    sieve_build(&pt, mapBytes, (mapBytes << 4) - 1, threads);
    td_primes(&pt, (uint64_t)smax * smax / 4 + 4);
    /* solve in parallel: one slot per listed sum */
    uint32_t *low = calloc(n ? n : 1, 4), *high = calloc(n ? n : 1, 4);
    _Atomic size_t next = 0;
    solve_job_t job = {&pt, sums, low, high, n, &next};
    pthread_t th[256];
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, solve_worker, &job);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    for (size_t i = 0; i < n; ++i) printf("%u %u %u\n", sums[i], low[i], high[i]);
    free(sums); free(low); free(high); free(pt.map); free(pt.tdPrimes);
    return 0;
}

int main(int argc, char **argv)
{
    int threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    uint64_t primemem = SEG_PRIMEMEM_DEFAULT, mapCap = 0;
    const char *dumpMap = NULL;
    int counters = 0, solve = 0, chunksOnly = 0, mrOnly = 0;
    uint32_t sample = 0;
    long long pos[2];
    int npos = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--primemem") && i + 1 < argc) primemem = strtoull(argv[++i], NULL, 10);
        else if (!strcmp(argv[i], "--map-cap") && i + 1 < argc) mapCap = strtoull(argv[++i], NULL, 10);
        else if (!strcmp(argv[i], "--dump-map") && i + 1 < argc) dumpMap = argv[++i];
        else if (!strcmp(argv[i], "--counters")) counters = 1;
        else if (!strcmp(argv[i], "--solve")) solve = 1;
        else if (!strcmp(argv[i], "--chunks")) chunksOnly = 1;
        else if (!strcmp(argv[i], "--mr")) mrOnly = 1;
        else if (!strcmp(argv[i], "--search-sample") && i + 1 < argc) sample = (uint32_t)atoi(argv[++i]);
        else if (argv[i][0] == '-' && argv[i][1] == '-') {
            fprintf(stderr, "ffref: unknown option %s\n", argv[i]);
            return 2;
        } else if (npos < 2) pos[npos++] = atoll(argv[i]);
    }
    if (threads < 1) threads = 1;
    if (solve) return run_solve(threads, mapCap);
    if (mrOnly) {
        unsigned long long n;
        while (scanf("%llu", &n) == 1) printf("%llu %d\n", n, (n > 17 && (n & 1)) ? miller_rabin(n) : -1);
        return 0;
    }
    /* positional handling of ff_seg main (:844-855) */
    uint32_t sumStart, sumLimit;
    if (npos >= 2) {
        sumStart = (uint32_t)(int)pos[0] | 1;
        sumLimit = (uint32_t)(int)pos[1];
    } else {
        sumStart = 5;
        sumLimit = npos == 1 ? (uint32_t)(int)pos[0] : 2627;
    }
    if (chunksOnly) {
        if (sumStart < 5) sumStart = 5;
        if (sumStart > sumLimit) sumLimit = sumStart;
        chunk_t ch[2 * SEG_PARTITION_THREADS];
        int n = seg_partition(sumStart, sumLimit, ch);
        for (int i = 0; i < n; ++i) printf("%u %u\n", ch[i].lo, ch[i].hi);
        return 0;
    }
    return run_seg(sumStart, sumLimit, primemem, threads, dumpMap, counters, sample);
}
