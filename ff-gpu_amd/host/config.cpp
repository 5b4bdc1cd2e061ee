#include "config.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>

namespace ffg {

static std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool parseSize(const std::string &s, uint64_t *out)
{
    if (s.empty() || lower(s) == "auto") return false;
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || v < 0.0) return false;
    std::string unit = lower(end);
    if (!unit.empty() && unit.back() == 'b') unit.pop_back();
    double mult;
    if (unit.empty() || unit == "byte") mult = 1.0;
    else if (unit == "k" || unit == "ki") mult = 1024.0;
    else if (unit == "m" || unit == "mi") mult = 1048576.0;
    else if (unit == "g" || unit == "gi") mult = 1073741824.0;
    else return false;
    const double bytes = v * mult;
    if (bytes > 18446744073709551615.0) return false;
    uint64_t r = (uint64_t)(bytes + 0.5);
    if (r == 0 && bytes > 0.0) r = 1;
    *out = r;
    return true;
}

bool parseFraction(const std::string &s, double *out)
{
    if (s.empty()) return false;
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || v < 0.0) return false;
    *out = v;
    return true;
}

static bool parsePairs(const std::string &spec, std::map<std::string, double> *out, bool lowerKey, bool positive)
{
    out->clear();
    size_t pos = 0;
    for (;;) {
        const size_t comma = spec.find(',', pos);
        const std::string item = spec.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
        const size_t eq = item.find('=');
        if (item.empty() || eq == std::string::npos) return false;
        std::string key = item.substr(0, eq);
        const std::string val = item.substr(eq + 1);
        if (key.empty() || val.empty()) return false;
        if (lowerKey) key = lower(key);
        double f = 0;
        if (!parseFraction(val, &f)) return false;
        if (positive && f <= 0.0) return false;
        (*out)[key] = f;
        if (comma == std::string::npos) return true;
        pos = comma + 1;
    }
}

bool parseDeviceFractionSpec(const std::string &spec, std::map<std::string, double> *out)
{
    return parsePairs(spec, out, false, false);
}

int loadEnv(Config *cfg)
{
    const char *v;
    uint64_t sz = 0;
    double d = 0;
    if ((v = std::getenv("FF_VRAM_FRACTION")) && *v) {
        if (!parseFraction(v, &d)) {
            std::fprintf(stderr, "[ff_sieve] error: malformed FF_VRAM_FRACTION='%s'\n", v);
            return -1;
        }
        cfg->globalFraction = d;
    }
    if ((v = std::getenv("FF_DEVICE_VRAM_FRACTION")) && *v) {
        if (!parseDeviceFractionSpec(v, &cfg->deviceFractions)) {
            std::fprintf(stderr, "[ff_sieve] error: malformed FF_DEVICE_VRAM_FRACTION='%s'\n", v);
            return -1;
        }
    }
    struct {
        const char *name;
        bool *flag;
        uint64_t *dst;
    } sizes[] = {{"FF_VRAM_BUDGET", &cfg->hasBudgetCap, &cfg->budgetCapBytes},
                 {"FF_SCRATCH", &cfg->hasScratch, &cfg->scratchBytes},
                 {"FF_SLAB_SIZE", nullptr, &cfg->slabSizeBytes}};
    for (auto &s : sizes) {
        if ((v = std::getenv(s.name)) && *v) {
            if (!parseSize(v, &sz)) {
                std::fprintf(stderr, "[ff_sieve] error: malformed %s='%s'\n", s.name, v);
                return -1;
            }
            if (s.flag) *s.flag = true;
            *s.dst = sz;
        }
    }
    if ((v = std::getenv("FF_HOST_TIER_CAP")) && *v) {
        if (lower(v) == "auto") cfg->hostTierAuto = true;
        else if (parseSize(v, &sz)) cfg->hostTierCapBytes = sz;
        else {
            std::fprintf(stderr, "[ff_sieve] error: malformed FF_HOST_TIER_CAP='%s'\n", v);
            return -1;
        }
    }
    if ((v = std::getenv("FF_PULL_WEIGHTS")) && *v) {
        if (!parsePairs(v, &cfg->pullWeights, true, true)) {
            std::fprintf(stderr, "[ff_sieve] error: malformed FF_PULL_WEIGHTS='%s'\n", v);
            return -1;
        }
    }
    if ((v = std::getenv("FF_DISABLE_DEVICE")) && *v) {
        const std::string b = lower(v);
        if (b != "amd" && b != "nvidia") {
            std::fprintf(stderr, "[ff_sieve] error: malformed FF_DISABLE_DEVICE='%s'\n", v);
            return -1;
        }
        cfg->disableVendor = b;
    }
    return 0;
}

int parseArgs(int argc, char **argv, Config *cfg, std::vector<std::string> *positionals)
{
    for (int i = 1; i < argc; ++i) {
        const std::string tok = argv[i];
        if (tok.rfind("--", 0) != 0) {
            positionals->push_back(tok);
            continue;
        }
        std::string name = tok, value;
        bool hasValue = false;
        const size_t eq = tok.find('=');
        if (eq != std::string::npos) {
            name = tok.substr(0, eq);
            value = tok.substr(eq + 1);
            hasValue = true;
        }
        auto need = [&](const char *flag) -> const char * {
            if (hasValue) return value.c_str();
            if (i + 1 < argc) return argv[++i];
            std::fprintf(stderr, "[ff_sieve] error: %s requires a value\n", flag);
            return nullptr;
        };
        auto size = [&](uint64_t *dst) -> bool {
            const char *v = need(name.c_str());
            if (!v) return false;
            if (!parseSize(v, dst)) {
                std::fprintf(stderr, "[ff_sieve] error: invalid size '%s'\n", v);
                return false;
            }
            return true;
        };
        if (name == "--list-devices") cfg->listDevices = true;
        else if (name == "--no-host-tier") cfg->noHostTier = true;
        else if (name == "--gpu-search") cfg->gpuSearch = true;
        else if (name == "--no-gpu") {
            std::fprintf(stderr, "[ff_sieve] error: --no-gpu: this build searches on the GPU only\n");
            return -1;
        } else if (name == "--vram-fraction") {
            const char *v = need("--vram-fraction");
            if (!v || !parseFraction(v, &cfg->globalFraction)) {
                if (v) std::fprintf(stderr, "[ff_sieve] error: invalid fraction '%s'\n", v);
                return -1;
            }
        } else if (name == "--device-vram-fraction") {
            const char *v = need("--device-vram-fraction");
            if (!v || !parseDeviceFractionSpec(v, &cfg->deviceFractions)) {
                if (v) std::fprintf(stderr, "[ff_sieve] error: invalid --device-vram-fraction '%s'\n", v);
                return -1;
            }
        } else if (name == "--vram-budget") {
            if (!size(&cfg->budgetCapBytes)) return -1;
            cfg->hasBudgetCap = true;
        } else if (name == "--scratch") {
            if (!size(&cfg->scratchBytes)) return -1;
            cfg->hasScratch = true;
        } else if (name == "--slab-size") {
            if (!size(&cfg->slabSizeBytes)) return -1;
        } else if (name == "--primemem") {
            if (!size(&cfg->primeMem)) return -1;
        } else if (name == "--map-cap") {
            const char *v = need("--map-cap");
            if (!v) return -1;
            char *end = nullptr;
            cfg->mapCap = std::strtoull(v, &end, 10);
            if (end == v || *end) {
                std::fprintf(stderr, "[ff_sieve] error: invalid --map-cap '%s'\n", v);
                return -1;
            }
        } else if (name == "--host-tier-cap") {
            const char *v = need("--host-tier-cap");
            if (!v) return -1;
            if (lower(v) == "auto") cfg->hostTierAuto = true;
            else if (!parseSize(v, &cfg->hostTierCapBytes)) {
                std::fprintf(stderr, "[ff_sieve] error: invalid size '%s'\n", v);
                return -1;
            }
        } else if (name == "--devices") {
            const char *v = need("--devices");
            if (!v) return -1;
            const std::string b = lower(v);
            if (b != "amd" && b != "nvidia") {
                std::fprintf(stderr, "[ff_sieve] error: invalid --devices backend '%s'\n", v);
                return -1;
            }
            cfg->deviceFilter = b;
        } else if (name == "--device") {
            const char *v = need("--device");
            if (!v) return -1;
            cfg->device = std::atoi(v);
        } else if (name == "--search-grab" || name == "--search-eager" || name == "--search-blocks-per-cu" ||
                   name == "--sieve-segment" || name == "--sieve-blocks-per-cu") {
            const char *v = need(name.c_str());
            if (!v) return -1;
            char *end = nullptr;
            const unsigned long x = std::strtoul(v, &end, 10);
            if (end == v || *end) {
                std::fprintf(stderr, "[ff_sieve] error: invalid %s '%s'\n", name.c_str(), v);
                return -1;
            }
            if (name == "--search-grab") cfg->searchGrab = (uint32_t)x;
            else if (name == "--search-eager") cfg->searchEager = (uint32_t)x;
            else if (name == "--search-blocks-per-cu") cfg->searchBlocksPerCU = (uint32_t)x;
            else if (name == "--sieve-segment") cfg->sieveSegmentBytes = (uint32_t)x;
            else cfg->sieveBlocksPerCU = (uint32_t)x;
        } else if (name == "--tune") {
            const char *v = need("--tune");
            if (!v) return -1;
            cfg->tuneFile = v;
        } else if (name == "--batch" || name == "--serve" || name == "--autotune") {
            cfg->mode = name.substr(2);
        } else if (name == "--opt") {
            const char *v = need("--opt");
            if (!v) return -1;
            cfg->modeArgs.push_back(v);
        } else if (name == "--dump-map") {
            const char *v = need("--dump-map");
            if (!v) return -1;
            cfg->dumpMapFile = v;
        } else {
            std::fprintf(stderr, "[ff_sieve] error: unknown option '%s'\n", name.c_str());
            return -1;
        }
    }
    return 0;
}

int loadTuneFile(Config *cfg, const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) {
        std::fprintf(stderr, "[ff_sieve] error: cannot read tuning file %s\n", path.c_str());
        return -1;
    }
    char buf[4096];
    const size_t n = std::fread(buf, 1, sizeof buf - 1, f);
    std::fclose(f);
    buf[n] = 0;
    struct {
        const char *key;
        uint32_t *dst;
    } keys[] = {{"search_grab", &cfg->searchGrab},
                {"search_eager_splits", &cfg->searchEager},
                {"search_blocks_per_cu", &cfg->searchBlocksPerCU},
                {"sieve_segment_bytes", &cfg->sieveSegmentBytes},
                {"sieve_blocks_per_cu", &cfg->sieveBlocksPerCU}};
    for (auto &k : keys) {
        const char *p = std::strstr(buf, k.key);
        if (!p) continue;
        p += std::strlen(k.key);
        while (*p && (*p == '"' || *p == ':' || *p == '=' || *p == ' ')) ++p;
        char *end = nullptr;
        const unsigned long v = std::strtoul(p, &end, 10);
        if (end == p) {
            std::fprintf(stderr, "[ff_sieve] error: tuning file %s: bad value for %s\n", path.c_str(), k.key);
            return -1;
        }
        if (*k.dst == 0) *k.dst = (uint32_t)v;  // explicit flags win
    }
    return 0;
}

bool slabSizeValueAligned(uint64_t slabSizeBytes)
{
    if (slabSizeBytes > UINT64_MAX / 8) return false;
    return (slabSizeBytes * 8) % 16 == 0;
}

int validateConfig(Config *cfg)
{
    if (cfg->globalFraction < 0.10 || cfg->globalFraction > 1.0) {
        std::fprintf(stderr, "[ff_sieve] validation error: --vram-fraction %.6g is outside [0.10, 1.0]\n",
                     cfg->globalFraction);
        return -1;
    }
    for (const auto &kv : cfg->deviceFractions)
        if (kv.second < 0.10 || kv.second > 1.0) {
            std::fprintf(stderr, "[ff_sieve] validation error: device fraction '%s=%.6g' is outside [0.10, 1.0]\n",
                         kv.first.c_str(), kv.second);
            return -1;
        }
    if (cfg->hasBudgetCap && cfg->budgetCapBytes == 0) {
        std::fprintf(stderr, "[ff_sieve] validation error: --vram-budget must be > 0\n");
        return -1;
    }
    if (cfg->hasScratch && cfg->scratchBytes == 0) {
        std::fprintf(stderr, "[ff_sieve] validation error: --scratch must be > 0\n");
        return -1;
    }
    if (cfg->slabSizeBytes == 0 || !slabSizeValueAligned(cfg->slabSizeBytes)) {
        std::fprintf(stderr, "[ff_sieve] validation error: --slab-size %llu B is zero or not 16-value aligned\n",
                     (unsigned long long)cfg->slabSizeBytes);
        return -1;
    }
    if (cfg->primeMem < 16) {
        std::fprintf(stderr, "[ff_sieve] validation error: --primemem must be >= 16\n");
        return -1;
    }
    if (!cfg->tuneFile.empty() && loadTuneFile(cfg, cfg->tuneFile) != 0) return -1;
    if (cfg->noHostTier) {
        cfg->hostTierCapBytes = 0;
        cfg->hostTierAuto = false;
    }
    if (cfg->hostTierAuto) {
        const long pages = sysconf(_SC_PHYS_PAGES), psz = sysconf(_SC_PAGE_SIZE);
        const uint64_t ram = (pages > 0 && psz > 0) ? (uint64_t)pages * (uint64_t)psz : 0;
        cfg->hostTierCapBytes = ram > (4ull << 30) ? ram - (4ull << 30) : 0;
        cfg->hostTierAuto = false;
    }
    return 0;
}

}  // namespace ffg
