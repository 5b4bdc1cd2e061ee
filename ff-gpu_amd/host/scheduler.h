// Slab scheduling of the sieve on one device.
//
// Single-device restatement of the pull scheduler's slab loop
// (source/pull_scheduler.cpp:149-219, :295-449): the map is sieved in slabs
// of --slab-size bytes on a compute stream; when the caller wants the map on
// the host (--dump-map), each finished slab is copied out on a second stream
// while the next slab computes, ordered by an event per slab (the
// reference enqueued its copy before the wait, :168-176, and re-copied
// everything afterwards; here the copy waits on the slab's event).
#ifndef FFG_SCHEDULER_H
#define FFG_SCHEDULER_H

#include <cstdint>

#include "ffx.h"

namespace ffg {

struct SieveTiming {
    double kernelMs = 0;   // first launch to last slab event (device clock)
    double wallMs = 0;     // host wall including copies
    uint64_t slabs = 0;
};

// Sieves map bytes [0, mapBytes) into d_map.  slabBytes == 0 (or >= mapBytes
// with no host copy) means one launch.  hostOut, when non-null, receives the
// whole map.
int sieveMap(const ffx_sieve_plan *plan, void *d_map, uint64_t mapBytes, uint64_t slabBytes, uint8_t *hostOut,
             void *stream, SieveTiming *timing);

// Host tier: sieves map bytes [0, mapBytes) into hostOut with no
// device-resident map, for maps larger than the device backing
// (source/pull_scheduler.cpp:164-200 computes such "homeless" slabs in a
// staging buffer).  Slab i is computed into device buffer i % ring on
// `stream` and copied into hostOut on a copy stream; the compute of slab
// i + ring waits on the device for slab i's copy, so the host thread only
// enqueues and compute overlaps the copies.  Device memory used:
// ring * slabBytes.  hostOut should come from ffx_alloc_pinned for the copies
// to run at the link's rate.
int sieveMapToHost(const ffx_sieve_plan *plan, uint64_t mapBytes, uint64_t slabBytes, int ring, uint8_t *hostOut,
                   void *stream, SieveTiming *timing);

}  // namespace ffg

#endif
