// Internal helpers shared by the libffx translation units (gfx950 only).
#ifndef FFX_INTERNAL_H
#define FFX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "ffx.h"

namespace ffx {

// Records a formatted message as the calling thread's last error and
// returns `code` (always negative).
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Maps a HIP status to an ffx code, recording the call site on failure.
int hip_fail(hipError_t e, const char *what);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Once per device, with the device current (ffx_init): takes the device
// address of the allocation-free search's scratch array and fills its
// tables (ffx_search.hip).
int noalloc_setup(int device);

}  // namespace ffx

#define FFX_HIP(call)                                                  \
    do {                                                               \
        hipError_t ffx_e_ = (call);                                    \
        if (ffx_e_ != hipSuccess) return ffx::hip_fail(ffx_e_, #call); \
    } while (0)

#endif  // FFX_INTERNAL_H
