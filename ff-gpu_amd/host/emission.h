// Byte-exact ff_seg stdout (reference/segmentedSieve.C:645-677, :816-819,
// :764-791, :922; source/m4/gpu_search_emission.cpp:34-97).
#ifndef FFG_EMISSION_H
#define FFG_EMISSION_H

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "ffx.h"
#include "partition.h"

namespace ffg {

// One solution line, numbered `ordinal`, newline-terminated.
std::string formatLine(uint64_t ordinal, const ffx_solution &s);

// Prints the solutions of every chunk in chunk order (sorted[] ascending by
// sum, one entry per solved sum), numbering lines from *ordinal+1 on.
// Returns bytes written.
uint64_t emitChunks(FILE *out, const std::vector<SumChunk> &chunks, const ffx_solution *sorted, size_t n,
                    uint64_t *ordinal);

// "Prime time: X μs\nFreudenthal time: Y μs\n"
void emitTimes(FILE *out, uint64_t primeUs, uint64_t searchUs);

}  // namespace ffg

#endif
