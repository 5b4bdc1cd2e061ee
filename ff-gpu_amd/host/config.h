// ff_sieve configuration: command-line flags and FF_* environment variables,
// the command line winning over the environment and both over the defaults.
//
// The flag and variable names and the syntax of their values are those of
// the reference's ff_sieve (source/config.h:15-85), so its command lines keep
// working; what they configure is one gfx950 device.  Knobs of the
// reference's multi-vendor scheduler are refused with a message instead of
// being stored.  New knobs: --primemem (ff_seg's PRIMEMEM, the header and map
// cap), --map-cap (largest value the device bitmap holds; Miller-Rabin
// answers above it), kernel tuning, and the batched/served/autotune modes.
#ifndef FFG_CONFIG_H
#define FFG_CONFIG_H

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace ffg {

struct Config {
    double globalFraction = 0.90;                   // --vram-fraction / FF_VRAM_FRACTION
    std::map<std::string, double> deviceFractions;  // --device-vram-fraction / FF_DEVICE_VRAM_FRACTION
    bool hasBudgetCap = false;                      // --vram-budget / FF_VRAM_BUDGET
    uint64_t budgetCapBytes = 0;
    bool hasScratch = false;                        // --scratch / FF_SCRATCH
    uint64_t scratchBytes = 0;
    uint64_t slabSizeBytes = 1ull << 30;            // --slab-size / FF_SLAB_SIZE
    bool hostTierAuto = false;                      // --host-tier-cap / FF_HOST_TIER_CAP ("auto" or a size)
    uint64_t hostTierCapBytes = 0;
    bool noHostTier = false;                        // --no-host-tier
    bool listDevices = false;                       // --list-devices
    std::string dumpMapFile;                        // --dump-map <file>
    uint64_t primeMem = 85ull << 30;                // --primemem
    uint64_t mapCap = 0;                            // --map-cap (0 = ff_seg's M)
    int device = 0;                                 // --device
    // kernel tuning (--tune FILE written by --autotune, or single flags)
    uint32_t searchGrab = 0, searchEager = 0, searchBlocksPerCU = 0;  // 0 = library default
    uint32_t sieveSegmentBytes = 0, sieveBlocksPerCU = 0;
    std::string tuneFile;
    // modes
    std::string mode;                               // "" (plain), "batch", "serve", "autotune"
    std::vector<std::string> modeArgs;              // key=value options of the mode
};

// Environment, then argv (non-option words go to *positionals), then the
// checks and derived values (tune file, host tier).  0, or -1 after a
// message on stderr.
int loadConfig(int argc, char **argv, Config *cfg, std::vector<std::string> *positionals);

// Value syntax shared with the reference's ff_sieve: a byte count with an
// optional K/M/G unit (KiB = KB = K = 1024 and so on, case-insensitive,
// fractional counts allowed: "1.5GiB"); a non-negative decimal fraction; a
// comma-separated list of key=fraction.
bool readByteSize(const std::string &text, uint64_t *out);
bool readFraction(const std::string &text, double *out);
bool readFractionMap(const std::string &text, std::map<std::string, double> *out);

// Reads the key=value lines (or a flat JSON object of numbers) of a tuning
// file into cfg's tuning fields that no flag set.  0, or -1 with a message.
int loadTuneFile(Config *cfg, const std::string &path);

}  // namespace ffg

#endif
