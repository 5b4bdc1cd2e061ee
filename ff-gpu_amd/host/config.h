// ff_sieve configuration: CLI flags and FF_* environment variables,
// precedence CLI > env > default (source/config.h:15-85, config.cpp:156-445).
//
// Kept knobs are those of one gfx950 device: VRAM fraction/budget, scratch,
// slab size, host tier (validated; the map never leaves HBM unless dumped),
// --dump-map, --list-devices.  Vendor selection (--devices, FF_DISABLE_DEVICE,
// FF_PULL_WEIGHTS) is parsed with the reference's error behaviour; only
// "amd" can match.  New knobs: --primemem (ff_seg's PRIMEMEM, the header and
// map cap), --map-cap (largest value the device bitmap holds; Miller-Rabin
// answers above it).
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
    bool hostTierAuto = false;                      // --host-tier-cap / FF_HOST_TIER_CAP
    uint64_t hostTierCapBytes = 0;
    bool noHostTier = false;                        // --no-host-tier
    bool listDevices = false;                       // --list-devices
    std::string deviceFilter;                       // --devices amd|nvidia
    std::string disableVendor;                      // FF_DISABLE_DEVICE
    std::map<std::string, double> pullWeights;      // FF_PULL_WEIGHTS
    std::string dumpMapFile;                        // --dump-map <file>
    bool gpuSearch = true;                          // --gpu-search (the default here)
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

// Reads the key=value lines (or a flat JSON object of numbers) of a tuning
// file into cfg's tuning fields.  Returns 0, or -1 with a message.
int loadTuneFile(Config *cfg, const std::string &path);

bool parseSize(const std::string &s, uint64_t *out);
bool parseFraction(const std::string &s, double *out);
bool parseDeviceFractionSpec(const std::string &spec, std::map<std::string, double> *out);
int loadEnv(Config *cfg);
int parseArgs(int argc, char **argv, Config *cfg, std::vector<std::string> *positionals);
int validateConfig(Config *cfg);
bool slabSizeValueAligned(uint64_t slabSizeBytes);

}  // namespace ffg

#endif
