/*
 * ffx.h -- C ABI of libffx, the gfx950 (MI355X) sieve + Freudenthal search.
 *
 * One backend, one architecture: every entry point runs HIP on the device
 * selected by ffx_init().  All types are POD; device buffers are `void*`,
 * streams and events are opaque `void*` handles.  Return value 0 means
 * success; a negative errno-style code means failure, and ffx_last_error()
 * then returns a message for the calling thread.  No entry point
 * synchronizes the device unless its name says so (sync / copy without
 * _async).  Buffers are owned by the caller except where a plan object says
 * it owns scratch.
 *
 * Each entry point cites the reference interface it replaces.  The reference
 * selects per-architecture function-pointer tables by vendor string
 * (source/sieve_slab_engine.h:49-101, source/m4/gpu_search_kernel.cpp:31-66,
 * source/devabstraction.h:70-126); libffx exports plain functions instead.
 */
#ifndef FFX_H
#define FFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFX_ABI_VERSION 1

/* Return codes (negative errno values). */
#define FFX_OK 0
#define FFX_EIO (-5)
#define FFX_ENOMEM (-12)
#define FFX_ENODEV (-19)
#define FFX_EINVAL (-22)
#define FFX_ERANGE (-34)

/* ---------------------------------------------------------------------- */
/* runtime                                                                 */
/* ---------------------------------------------------------------------- */

/* Device descriptor; the fields of ff::DeviceInfo (source/device_info.h:15-35)
 * that a single-vendor build still needs. */
typedef struct ffx_device_info {
    char vendor[8];        /* always "amd" */
    char name[256];
    char arch[64];         /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    char busId[32];        /* "%04x:%02x:%02x" */
    uint64_t freeBytes;
    uint64_t totalBytes;
    uint64_t sharedMemPerBlock;
    uint64_t sharedMemPerMultiprocessor;
    uint64_t l2Bytes;
    int32_t multiProcessorCount;
    int32_t maxThreadsPerBlock;
    int32_t warpSize;
    int32_t clockKHz;
    int32_t memoryClockKHz;
    int32_t memoryBusWidth;
    int32_t runtimeIndex;
    int32_t reserved;
} ffx_device_info;

/* ABI version the library was built with (FFX_ABI_VERSION). */
int ffx_abi_version(void);
/* Replaces ffdev::DevInit (source/devabstraction.h:70-88): selects `device`
 * for the calling thread; idempotent. */
int ffx_init(int device);
/* Message of the last failure on the calling thread ("" if none). */
const char *ffx_last_error(void);
/* Replaces ff_enum_hip (source/hip_enum.cpp:32). */
int ffx_device_count(int *count);
int ffx_device_info_get(int device, ffx_device_info *out);
/* hipMemGetInfo of the current device. */
int ffx_mem_info(uint64_t *freeBytes, uint64_t *totalBytes);

/* Memory and copies: SievePoolOps alloc/memset/free/copyH2D/copyD2H
 * (source/sieve_slab_engine.h:49-62) and ffdev::DevAlloc/DevFree/DevCopy
 * (source/devabstraction.h:89-93).  ffx_alloc returns NULL on failure. */
void *ffx_alloc(uint64_t bytes);
int ffx_free(void *d);
int ffx_memset_async(void *d, int value, uint64_t bytes, void *stream);
int ffx_copy_h2d(void *d, const void *h, uint64_t bytes);
int ffx_copy_d2h(void *h, const void *d, uint64_t bytes);
int ffx_copy_h2d_async(void *d, const void *h, uint64_t bytes, void *stream);
int ffx_copy_d2h_async(void *h, const void *d, uint64_t bytes, void *stream);
/* Pinned host memory: SievePoolOps allocPinned/freePinned (:86-88). */
void *ffx_alloc_pinned(uint64_t bytes);
int ffx_free_pinned(void *h);

/* Streams and events: SievePoolOps createStream .. elapsedEvents (:70-85).
 * Streams are created at the default priority. */
void *ffx_stream_create(void);
int ffx_stream_destroy(void *stream);
int ffx_stream_sync(void *stream);
void *ffx_event_create(void);
int ffx_event_destroy(void *ev);
int ffx_event_record(void *ev, void *stream);
int ffx_event_wait(void *ev, void *stream);
int ffx_event_sync(void *ev);
/* Non-blocking: 0 when the event has completed, 1 when not yet, < 0 on
 * error (hipEventQuery). */
int ffx_event_query(void *ev);
int ffx_event_elapsed_ms(void *start, void *end, double *ms);
int ffx_device_sync(void);

/* HIP graphs (no reference counterpart): capture the launches submitted to
 * a created stream, then replay them with one call.  Captured launches read
 * their arguments as they were at capture time. */
typedef struct ffx_graph ffx_graph;
int ffx_graph_begin(void *stream);
int ffx_graph_end(void *stream, ffx_graph **out);
int ffx_graph_launch(ffx_graph *graph, void *stream);
int ffx_graph_destroy(ffx_graph *graph);

/* ---------------------------------------------------------------------- */
/* sieve                                                                    */
/*                                                                          */
/* The map is ff_seg's odd-only bitmap: the bit of odd value n lives in     */
/* byte n>>4 under mask 0x80>>((n>>1)&7).  For odd n <= maxPrimeMapValue   */
/* the bit is 1 iff n is prime; bits of odd values above it are 1 (never   */
/* struck), as in Prime::SegmentFill (reference/segmentedSieve.C:255-270).  */
/* ---------------------------------------------------------------------- */

typedef struct ffx_sieve_plan ffx_sieve_plan;

/* Replaces SieveEngine::prepare + the prime upload of SievePoolOps::
 * slabCompute (source/sieve_engine.cpp:114-133, source/sieve_slab_engine.cpp:
 * 279-312): generates the sieving primes 3..isqrt(maxPrimeMapValue) on the
 * host, builds the small-prime word patterns and uploads both. */
int ffx_sieve_plan_create(ffx_sieve_plan **out, uint64_t maxPrimeMapValue);
/* Same, with the caller's odd sieving primes (the h_primes/numPrimes pair of
 * SievePoolOps::slabCompute).  The list must hold every odd prime up to
 * isqrt(maxPrimeMapValue), ascending; a 2 at the front is ignored. */
int ffx_sieve_plan_create_from_primes(ffx_sieve_plan **out, const uint32_t *h_primes, uint32_t numPrimes,
                                      uint64_t maxPrimeMapValue);
int ffx_sieve_plan_destroy(ffx_sieve_plan *plan);

/* Tunable sieve parameters (no reference counterpart; the reference fixes
 * its kernel geometry at compile time, source/sieve_slab_kernel.h:40-67). */
#define FFX_SIEVE_SEGMENT_BYTES 1  /* LDS tile: map bytes per segment, 32768 or 114688 */
#define FFX_SIEVE_BLOCKS_PER_CU 2  /* persistent grid: 1024-thread workgroups per CU (1..8) */
#define FFX_SIEVE_LANE_CARRY 3     /* at most this many one-per-lane primes per thread keep their next
                                      strike across segments (0..64, default 64; the kernel holds
                                      at most 12 with 32 KiB segments, 32 or 64 with 112 KiB) */
#define FFX_SIEVE_HOLD_PRIMES 4    /* 112 KiB segments: carried primes held in registers (1) or
                                      re-read each segment (2); 0 (default) holds them */
int ffx_sieve_plan_set(ffx_sieve_plan *plan, int param, uint32_t value);
int ffx_sieve_plan_get(const ffx_sieve_plan *plan, int param, uint32_t *value);

/* Replaces SievePoolOps::slabComputeAsync (source/sieve_slab_engine.h:93-97):
 * writes the map bytes covering values [segLo, segHi) into d_slab, byte i of
 * d_slab being map byte segLo/16 + i.  segLo must be a multiple of 16; the
 * slab holds ceil((segHi-segLo)/16) bytes.  Priming (0xFF fill, clearing
 * value 1) is fused into the kernel, so no memset is needed first. */
int ffx_sieve_slab_async(const ffx_sieve_plan *plan, uint64_t segLo, uint64_t segHi, void *d_slab, void *stream);
/* Whole map in one launch: bytes [0, mapBytes) of d_map. */
int ffx_sieve_map_async(const ffx_sieve_plan *plan, void *d_map, uint64_t mapBytes, void *stream);
/* Map bytes [byteLo, byteHi) of a resident map d_map (growth of a map that
 * is already partly sieved; bytes outside the range are not written). */
int ffx_sieve_map_range_async(const ffx_sieve_plan *plan, void *d_map, uint64_t byteLo, uint64_t byteHi,
                              void *stream);

/* ---------------------------------------------------------------------- */
/* search                                                                   */
/*                                                                          */
/* F(s) for odd s: the statement-1 filter and DoesSammyKnow of              */
/* reference/segmentedSieve.C:805-815, :529-639.  Primality is the bitmap   */
/* for n <= maxPrimeMapValue and deterministic Miller-Rabin above it        */
/* (IsPrime :283-296), so any map size gives the same answers.              */
/* ---------------------------------------------------------------------- */

/* Slot encoding of one odd sum: 0 = the sum is not emitted; otherwise
 * bits 0..27 hold the low term and the flags below tag both terms.  The high
 * term is sum - low. */
#define FFX_SLOT_LOW_MASK 0x0FFFFFFFu
#define FFX_TAG_LOW_PRIME 0x10000000u
#define FFX_TAG_LOW_POW2 0x20000000u
#define FFX_TAG_HIGH_PRIME 0x40000000u
#define FFX_TAG_HIGH_POW2 0x80000000u
/* Largest sum the search accepts: terms fit the slot encoding and every
 * product q <= s^2/16 keeps prime exponents below 32. */
#define FFX_MAX_SUM 0x03FFFFFFu

/* One emitted line, in ascending sum order after compaction. */
typedef struct ffx_solution {
    uint32_t sum;
    uint32_t low;
    uint32_t high;
    uint32_t tags; /* FFX_TAG_* bits */
} ffx_solution;

/* The reference's sum-indexed record (source/m4/gpu_search_kernel.h:37-42;
 * 32 bytes with its padding).  A zero record is an unsolved slot. */
typedef struct ffx_gpu_record {
    uint32_t sum;
    uint32_t pad0;
    uint64_t low;
    uint64_t high;
    uint8_t tag;
    uint8_t pad1[7];
} ffx_gpu_record;

typedef struct ffx_search_plan ffx_search_plan;

/* One interval of a batched search: the odd sums first, first+2, ..., last
 * fill slots slotOffset, slotOffset+1, ... of the slot array. */
typedef struct ffx_interval {
    uint32_t first;
    uint32_t last;
    uint64_t slotOffset;
} ffx_interval;

/* Tunable search parameters (ffx_search_plan_set / _get). */
#define FFX_SEARCH_GRAB 1           /* sums a wavefront takes per atomic (1..64) */
#define FFX_SEARCH_EAGER_SPLITS 2   /* splits per sum resolved eagerly before batching candidates */
#define FFX_SEARCH_BLOCKS_PER_CU 3  /* persistent grid: workgroups of 4 wavefronts per CU (1..16) */

/* Prepares the search for sums up to sumLimitMax: uploads the smallest-
 * prime-factor table used to factor terms (< sumLimitMax) and allocates the
 * plan's counters and compaction scratch. */
int ffx_search_plan_create(ffx_search_plan **out, uint32_t sumLimitMax);
int ffx_search_plan_destroy(ffx_search_plan *plan);

/* Per-sum slots: d_slots[(s - first)/2] for odd s in [first, sumLimit],
 * first = max(5, sumStart|1).  d_map holds at least (maxPrimeMapValue>>4)+1
 * bytes.  Work is pulled by persistent wavefronts; calls on one plan must be
 * ordered (one stream) because they share its work counter. */
int ffx_search_slots_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue, uint32_t sumStart,
                           uint32_t sumLimit, uint32_t *d_slots, void *stream);

/* Batched search (no reference counterpart; the reference launches one
 * range per call): every odd sum of nintervals disjoint intervals (device
 * array, slotOffset ascending, sums ascending) in ONE persistent launch;
 * wavefronts pull sums from a shared counter, largest first, so sub-ranges
 * of very different cost balance across the device.  lastSum is the
 * largest sum of any interval (checked against the plan). */
int ffx_search_intervals_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue,
                               const ffx_interval *d_intervals, uint32_t nintervals, uint64_t totalSlots,
                               uint32_t lastSum, uint32_t *d_slots, void *stream);

/* Same, but the slot of sum s is d_slots[(s - slotSumBase) / 2] whatever
 * interval holds it: one resident slot array for all sums of a stream of
 * batches (slotSumBase odd, <= every interval's first sum).  Slots at or
 * beyond slotCapacity are never written. */
int ffx_search_intervals_by_sum_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue,
                                      const ffx_interval *d_intervals, uint32_t nintervals, uint64_t totalSlots,
                                      uint32_t lastSum, uint32_t slotSumBase, uint32_t *d_slots, uint64_t slotCapacity,
                                      void *stream);

/* Tuning knobs of a plan (FFX_SEARCH_*); values out of range fail. */
int ffx_search_plan_set(ffx_search_plan *plan, int param, uint32_t value);
int ffx_search_plan_get(const ffx_search_plan *plan, int param, uint32_t *value);

/* Ordered compaction of slots into solutions: d_out[0..*d_count) in slot
 * order.  d_out must hold one entry per solved slot (nslots is an upper
 * bound).  Uses the plan's scratch. */
int ffx_compact_async(ffx_search_plan *plan, const uint32_t *d_slots, uint64_t nslots, uint32_t firstSum,
                      ffx_solution *d_out, uint32_t *d_count, void *stream);

/* Slots + compaction in one call; d_slots may be NULL (plan scratch). */
int ffx_search_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue, uint32_t sumStart,
                     uint32_t sumLimit, ffx_solution *d_out, uint32_t *d_count, void *stream);

/* Drop-in for SearchKernelRun_<arch> (source/m4/gpu_search_kernel.cpp:31-50,
 * typedef SearchKernelRunFn :59-61): fills the caller's zeroed 32-byte
 * sum-indexed records and adds the number of solutions to *d_count.
 * Synchronous (the reference relied on a later blocking copy). */
int ffx_search_records(int device, const uint8_t *d_map, uint64_t maxPrimeMapValue, uint64_t sumStart,
                       uint64_t sumLimit, uint32_t *d_count, ffx_gpu_record *d_records);

/* Largest sum ffx_search_records_noalloc accepts: its scratch array holds
 * the slots and factor table of odd sums up to here (6.6 MB per device). */
#define FFX_NOALLOC_MAX_SUM 0x001FFFFFu

/* No counterpart in the reference: ffx_search_records for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
 * allocating routine: its work counter, tables and slots live in a
 * fixed-size device array (one per device, set up by ffx_init(device), which
 * must have been called).  Same contract as ffx_search_records (the caller
 * zeroes d_records and *d_count; solved sums fill their records and add to
 * *d_count), except that it is asynchronous on `stream` and sumLimit is at
 * most FFX_NOALLOC_MAX_SUM (FFX_ERANGE above).  Calls that overlap on one
 * device race on the shared array: order them (one stream, or events). */
int ffx_search_records_noalloc(int device, const uint8_t *d_map, uint64_t maxPrimeMapValue, uint32_t sumStart,
                               uint32_t sumLimit, uint32_t *d_count, ffx_gpu_record *d_records, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FFX_H */
