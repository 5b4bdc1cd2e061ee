/* Diagnostic counters of the search kernel, exported only by
 * lib/libffx_counters.so (make -C ff-gpu_amd counters), which is libffx
 * built with FFX_PROBE_COUNTERS.  No counterpart in the reference.
 *
 * ffx_probe_counters copies the 16 running totals since the last reset into
 * out and, when reset != 0, clears them.  Indices:
 *   0 sums that reached the split enumeration   5 map reads of o (Power2Even)
 *   1 splits visited                            6 AllButOne checks
 *   2 splits past the residue filter            7 map reads inside AllButOne
 *   3 map reads of 2^k + d*o - 2                8 Miller-Rabin tests
 *   4 map reads of d + o*2^k - 2                9 splits listed
 * Returns 0, or a negative FFX_E* code with ffx_last_error(). */
#ifndef FFX_COUNTERS_H
#define FFX_COUNTERS_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int ffx_probe_counters(uint64_t *out, int reset);
#ifdef __cplusplus
}
#endif
#endif
