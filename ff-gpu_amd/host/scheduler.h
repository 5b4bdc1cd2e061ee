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

}  // namespace ffg

#endif
