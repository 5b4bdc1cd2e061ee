// libffx runtime: device selection, memory, copies, streams and events.
//
// One gfx950 HIP path replaces the reference's dual CUDA/ROCm device layer
// (source/devabstraction.cpp, source/hip_devabstraction.cpp:90-290) and the
// memory/stream/event members of SievePoolOps
// (source/sieve_slab_engine.h:49-98, source/sieve_slab_engine.cpp:240-443).
#include <cstring>

#include "ffx_internal.h"

namespace ffx {

static thread_local char g_err[512];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    int code = e == hipErrorOutOfMemory ? FFX_ENOMEM
               : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? FFX_ENODEV
               : e == hipErrorInvalidValue ? FFX_EINVAL
                                            : FFX_EIO;
    return fail(code, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace ffx

using namespace ffx;

extern "C" {

int ffx_abi_version(void) { return FFX_ABI_VERSION; }

const char *ffx_last_error(void) { return g_err; }

int ffx_init(int device)
{
    int n = 0;
    FFX_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(FFX_ENODEV, "ffx_init: device %d of %d", device, n);
    FFX_HIP(hipSetDevice(device));
    FFX_HIP(hipFree(nullptr));  // create the context now, not inside a timed region
    return noalloc_setup(device);
}

int ffx_device_count(int *count)
{
    if (!count) return fail(FFX_EINVAL, "ffx_device_count: null");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return FFX_OK;
    }
    FFX_HIP(e);
    return FFX_OK;
}

int ffx_device_info_get(int device, ffx_device_info *out)
{
    if (!out) return fail(FFX_EINVAL, "ffx_device_info_get: null");
    memset(out, 0, sizeof *out);
    hipDeviceProp_t p;
    FFX_HIP(hipGetDeviceProperties(&p, device));
    snprintf(out->vendor, sizeof out->vendor, "amd");
    snprintf(out->name, sizeof out->name, "%s", p.name);
    snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    snprintf(out->busId, sizeof out->busId, "%04x:%02x:%02x", p.pciDomainID, p.pciBusID, p.pciDeviceID);
    int prev = 0;
    FFX_HIP(hipGetDevice(&prev));
    FFX_HIP(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    hipError_t e = hipMemGetInfo(&fr, &tot);
    FFX_HIP(hipSetDevice(prev));
    FFX_HIP(e);
    out->freeBytes = fr;
    out->totalBytes = tot;
    out->sharedMemPerBlock = p.sharedMemPerBlock;
    out->sharedMemPerMultiprocessor = p.maxSharedMemoryPerMultiProcessor;
    out->l2Bytes = (uint64_t)p.l2CacheSize;
    out->multiProcessorCount = p.multiProcessorCount;
    out->maxThreadsPerBlock = p.maxThreadsPerBlock;
    out->warpSize = p.warpSize;
    out->clockKHz = p.clockRate;
    out->memoryClockKHz = p.memoryClockRate;
    out->memoryBusWidth = p.memoryBusWidth;
    out->runtimeIndex = device;
    return FFX_OK;
}

int ffx_mem_info(uint64_t *freeBytes, uint64_t *totalBytes)
{
    size_t fr = 0, tot = 0;
    FFX_HIP(hipMemGetInfo(&fr, &tot));
    if (freeBytes) *freeBytes = fr;
    if (totalBytes) *totalBytes = tot;
    return FFX_OK;
}

void *ffx_alloc(uint64_t bytes)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        hip_fail(e, "hipMalloc");
        return nullptr;
    }
    return p;
}

int ffx_free(void *d)
{
    if (d) FFX_HIP(hipFree(d));
    return FFX_OK;
}

int ffx_memset_async(void *d, int value, uint64_t bytes, void *stream)
{
    FFX_HIP(hipMemsetAsync(d, value, bytes, as_stream(stream)));
    return FFX_OK;
}

int ffx_copy_h2d(void *d, const void *h, uint64_t bytes)
{
    FFX_HIP(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return FFX_OK;
}

int ffx_copy_d2h(void *h, const void *d, uint64_t bytes)
{
    FFX_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
    return FFX_OK;
}

int ffx_copy_h2d_async(void *d, const void *h, uint64_t bytes, void *stream)
{
    FFX_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    return FFX_OK;
}

int ffx_copy_d2h_async(void *h, const void *d, uint64_t bytes, void *stream)
{
    FFX_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    return FFX_OK;
}

void *ffx_alloc_pinned(uint64_t bytes)
{
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        hip_fail(e, "hipHostMalloc");
        return nullptr;
    }
    return p;
}

int ffx_free_pinned(void *h)
{
    if (h) FFX_HIP(hipHostFree(h));
    return FFX_OK;
}

void *ffx_stream_create(void)
{
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        hip_fail(e, "hipStreamCreateWithFlags");
        return nullptr;
    }
    return s;
}

int ffx_stream_destroy(void *stream)
{
    if (stream) FFX_HIP(hipStreamDestroy(as_stream(stream)));
    return FFX_OK;
}

int ffx_stream_sync(void *stream)
{
    FFX_HIP(hipStreamSynchronize(as_stream(stream)));
    return FFX_OK;
}

void *ffx_event_create(void)
{
    hipEvent_t ev = nullptr;
    hipError_t e = hipEventCreate(&ev);
    if (e != hipSuccess) {
        hip_fail(e, "hipEventCreate");
        return nullptr;
    }
    return ev;
}

int ffx_event_destroy(void *ev)
{
    if (ev) FFX_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
    return FFX_OK;
}

int ffx_event_record(void *ev, void *stream)
{
    FFX_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), as_stream(stream)));
    return FFX_OK;
}

int ffx_event_wait(void *ev, void *stream)
{
    FFX_HIP(hipStreamWaitEvent(as_stream(stream), reinterpret_cast<hipEvent_t>(ev), 0));
    return FFX_OK;
}

int ffx_event_query(void *ev)
{
    const hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(ev));
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return hip_fail(e, "hipEventQuery");
}

int ffx_event_sync(void *ev)
{
    FFX_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(ev)));
    return FFX_OK;
}

int ffx_event_elapsed_ms(void *start, void *end, double *ms)
{
    float f = 0.f;
    FFX_HIP(hipEventElapsedTime(&f, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(end)));
    if (ms) *ms = f;
    return FFX_OK;
}

int ffx_device_sync(void)
{
    FFX_HIP(hipDeviceSynchronize());
    return FFX_OK;
}

}  // extern "C"

struct ffx_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" {

int ffx_graph_begin(void *stream)
{
    if (!stream) return fail(FFX_EINVAL, "graph capture needs a created stream");
    FFX_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
    return FFX_OK;
}

int ffx_graph_end(void *stream, ffx_graph **out)
{
    if (!out) return fail(FFX_EINVAL, "graph: null out");
    *out = nullptr;
    hipGraph_t g = nullptr;
    FFX_HIP(hipStreamEndCapture(as_stream(stream), &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate");
    }
    ffx_graph *r = new ffx_graph{g, ex};
    *out = r;
    return FFX_OK;
}

int ffx_graph_launch(ffx_graph *g, void *stream)
{
    if (!g) return fail(FFX_EINVAL, "graph: null");
    FFX_HIP(hipGraphLaunch(g->exec, as_stream(stream)));
    return FFX_OK;
}

int ffx_graph_destroy(ffx_graph *g)
{
    if (!g) return FFX_OK;
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
    return FFX_OK;
}

}  // extern "C"
