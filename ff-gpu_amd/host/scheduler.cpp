#include "scheduler.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

namespace ffg {

namespace {
struct Events {
    std::vector<void *> ev;
    ~Events()
    {
        for (void *e : ev) ffx_event_destroy(e);
    }
    void *make()
    {
        void *e = ffx_event_create();
        if (e) ev.push_back(e);
        return e;
    }
};
}  // namespace

int sieveMap(const ffx_sieve_plan *plan, void *d_map, uint64_t mapBytes, uint64_t slabBytes, uint8_t *hostOut,
             void *stream, SieveTiming *timing)
{
    const auto t0 = std::chrono::steady_clock::now();
    Events evs;
    void *start = evs.make(), *stop = evs.make();
    if (!start || !stop) return FFX_EIO;
    int rc = ffx_event_record(start, stream);
    if (rc) return rc;
    if (slabBytes == 0 || slabBytes > mapBytes) slabBytes = mapBytes;
    uint64_t nslabs = 0;
    if (!hostOut) {
        rc = ffx_sieve_map_async(plan, d_map, mapBytes, stream);
        nslabs = 1;
        if (rc) return rc;
        if ((rc = ffx_event_record(stop, stream))) return rc;
        if ((rc = ffx_event_sync(stop))) return rc;
    } else {
        // compute slab i on `stream`; copy slab i-1 through pinned staging on
        // `copy` after waiting for slab i-1's event
        void *copy = ffx_stream_create();
        if (!copy) return FFX_EIO;
        uint8_t *stage[2] = {(uint8_t *)ffx_alloc_pinned(slabBytes), (uint8_t *)ffx_alloc_pinned(slabBytes)};
        void *copied[2] = {evs.make(), evs.make()};
        bool pending[2] = {false, false};
        uint64_t pendLo[2] = {0, 0}, pendLen[2] = {0, 0};
        if (!stage[0] || !stage[1] || !copied[0] || !copied[1]) rc = FFX_ENOMEM;
        for (uint64_t lo = 0; !rc && lo < mapBytes; lo += slabBytes, ++nslabs) {
            const uint64_t len = mapBytes - lo < slabBytes ? mapBytes - lo : slabBytes;
            const int b = (int)(nslabs & 1);
            // staging buffer b is free once its previous copy has landed
            if (pending[b]) {
                if ((rc = ffx_event_sync(copied[b]))) break;
                std::memcpy(hostOut + pendLo[b], stage[b], pendLen[b]);
                pending[b] = false;
            }
            if ((rc = ffx_sieve_map_range_async(plan, d_map, lo, lo + len, stream))) break;
            void *done = evs.make();
            if (!done) { rc = FFX_EIO; break; }
            if ((rc = ffx_event_record(done, stream))) break;
            if ((rc = ffx_event_wait(done, copy))) break;
            if ((rc = ffx_copy_d2h_async(stage[b], (uint8_t *)d_map + lo, len, copy))) break;
            if ((rc = ffx_event_record(copied[b], copy))) break;
            pending[b] = true;
            pendLo[b] = lo;
            pendLen[b] = len;
        }
        if (!rc) rc = ffx_event_record(stop, stream);
        for (int b = 0; b < 2 && !rc; ++b)
            if (pending[b]) {
                rc = ffx_event_sync(copied[b]);
                if (!rc) std::memcpy(hostOut + pendLo[b], stage[b], pendLen[b]);
            }
        if (!rc) rc = ffx_event_sync(stop);
        ffx_free_pinned(stage[0]);
        ffx_free_pinned(stage[1]);
        ffx_stream_destroy(copy);
        if (rc) return rc;
    }
    double ms = 0;
    if ((rc = ffx_event_elapsed_ms(start, stop, &ms))) return rc;
    if (timing) {
        timing->kernelMs = ms;
        timing->slabs = nslabs;
        timing->wallMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return FFX_OK;
}

int sieveMapToHost(const ffx_sieve_plan *plan, uint64_t mapBytes, uint64_t slabBytes, int ring, uint8_t *hostOut,
                   void *stream, SieveTiming *timing)
{
    if (!plan || !hostOut || !stream || ring < 1 || ring > 8 || slabBytes == 0) return FFX_EINVAL;
    const auto t0 = std::chrono::steady_clock::now();
    Events evs;
    void *start = evs.make(), *stop = evs.make();
    void *copy = ffx_stream_create();
    std::vector<void *> buf((size_t)ring, nullptr), computed((size_t)ring, nullptr), copied((size_t)ring, nullptr);
    int rc = start && stop && copy ? FFX_OK : FFX_EIO;
    for (int r = 0; !rc && r < ring; ++r) {
        buf[(size_t)r] = ffx_alloc(slabBytes);
        computed[(size_t)r] = evs.make();
        copied[(size_t)r] = evs.make();
        if (!buf[(size_t)r]) rc = FFX_ENOMEM;
        else if (!computed[(size_t)r] || !copied[(size_t)r]) rc = FFX_EIO;
    }
    uint64_t nslabs = 0;
    if (!rc) rc = ffx_event_record(start, stream);
    for (uint64_t lo = 0; !rc && lo < mapBytes; lo += slabBytes, ++nslabs) {
        const uint64_t len = mapBytes - lo < slabBytes ? mapBytes - lo : slabBytes;
        const size_t r = (size_t)(nslabs % (uint64_t)ring);
        // the buffer is free once the copy of the slab it held has landed
        if (nslabs >= (uint64_t)ring && (rc = ffx_event_wait(copied[r], stream))) break;
        if ((rc = ffx_sieve_slab_async(plan, lo << 4, (lo + len) << 4, buf[r], stream))) break;
        if ((rc = ffx_event_record(computed[r], stream))) break;
        if ((rc = ffx_event_wait(computed[r], copy))) break;
        if ((rc = ffx_copy_d2h_async(hostOut + lo, buf[r], len, copy))) break;
        if ((rc = ffx_event_record(copied[r], copy))) break;
    }
    if (!rc) rc = ffx_event_record(stop, copy);
    if (!rc) rc = ffx_event_sync(stop);
    // drain both streams before the buffers go, also after an error
    if (copy) ffx_stream_sync(copy);
    ffx_stream_sync(stream);
    for (void *b : buf) ffx_free(b);
    ffx_stream_destroy(copy);
    if (rc) return rc;
    double ms = 0;
    if ((rc = ffx_event_elapsed_ms(start, stop, &ms))) return rc;
    if (timing) {
        timing->kernelMs = ms;
        timing->slabs = nslabs;
        timing->wallMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return FFX_OK;
}

}  // namespace ffg
