// ff_sieve configuration: one table of knobs drives the environment, the
// command line and their error messages; checkConfig then tests the whole.
#include "config.h"

#include <unistd.h>

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ffg {

namespace {

bool sameNoCase(const std::string &a, const char *b)
{
    if (a.size() != std::strlen(b)) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

bool readUnsigned(const std::string &text, uint64_t *out)
{
    if (text.empty() || !std::isdigit((unsigned char)text[0])) return false;
    char *end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(text.c_str(), &end, 10);
    if (*end || errno) return false;
    *out = v;
    return true;
}

// Each knob: its flag and/or environment name, whether the flag takes a
// value, and how a value is stored ("" when accepted, else why not).
using Apply = std::string (*)(Config &, const std::string &);

struct Knob {
    const char *flag;
    const char *env;
    bool takesValue;
    Apply apply;
};

std::string storeBytes(const std::string &v, uint64_t *dst, bool *given)
{
    uint64_t n = 0;
    if (!readByteSize(v, &n)) return "not a byte size (e.g. 8GiB, 512MiB, 4096)";
    *dst = n;
    if (given) *given = true;
    return "";
}

std::string storeCount(const std::string &v, uint32_t *dst)
{
    uint64_t n = 0;
    if (!readUnsigned(v, &n) || n > 0xFFFFFFFFull) return "not a non-negative integer";
    *dst = (uint32_t)n;
    return "";
}

const Knob kKnobs[] = {
    {"--vram-fraction", "FF_VRAM_FRACTION", true,
     [](Config &c, const std::string &v) -> std::string {
         return readFraction(v, &c.globalFraction) ? "" : "not a non-negative fraction";
     }},
    {"--device-vram-fraction", "FF_DEVICE_VRAM_FRACTION", true,
     [](Config &c, const std::string &v) -> std::string {
         return readFractionMap(v, &c.deviceFractions) ? "" : "not a list of device=fraction";
     }},
    {"--vram-budget", "FF_VRAM_BUDGET", true,
     [](Config &c, const std::string &v) { return storeBytes(v, &c.budgetCapBytes, &c.hasBudgetCap); }},
    {"--scratch", "FF_SCRATCH", true,
     [](Config &c, const std::string &v) { return storeBytes(v, &c.scratchBytes, &c.hasScratch); }},
    {"--slab-size", "FF_SLAB_SIZE", true,
     [](Config &c, const std::string &v) { return storeBytes(v, &c.slabSizeBytes, nullptr); }},
    {"--host-tier-cap", "FF_HOST_TIER_CAP", true,
     [](Config &c, const std::string &v) -> std::string {
         c.hostTierAuto = sameNoCase(v, "auto");
         return c.hostTierAuto ? "" : storeBytes(v, &c.hostTierCapBytes, nullptr);
     }},
    {"--no-host-tier", nullptr, false,
     [](Config &c, const std::string &) -> std::string {
         c.noHostTier = true;
         return "";
     }},
    {"--list-devices", nullptr, false,
     [](Config &c, const std::string &) -> std::string {
         c.listDevices = true;
         return "";
     }},
    {"--dump-map", nullptr, true,
     [](Config &c, const std::string &v) -> std::string {
         c.dumpMapFile = v;
         return v.empty() ? "an empty file name" : "";
     }},
    {"--primemem", nullptr, true, [](Config &c, const std::string &v) { return storeBytes(v, &c.primeMem, nullptr); }},
    {"--map-cap", nullptr, true,
     [](Config &c, const std::string &v) -> std::string {
         return readUnsigned(v, &c.mapCap) ? "" : "not a non-negative integer (a map value)";
     }},
    {"--device", nullptr, true,
     [](Config &c, const std::string &v) -> std::string {
         uint64_t d = 0;
         if (!readUnsigned(v, &d) || d > 1023) return "not a device index";
         c.device = (int)d;
         return "";
     }},
    {"--search-grab", nullptr, true, [](Config &c, const std::string &v) { return storeCount(v, &c.searchGrab); }},
    {"--search-eager", nullptr, true, [](Config &c, const std::string &v) { return storeCount(v, &c.searchEager); }},
    {"--search-blocks-per-cu", nullptr, true,
     [](Config &c, const std::string &v) { return storeCount(v, &c.searchBlocksPerCU); }},
    {"--sieve-segment", nullptr, true,
     [](Config &c, const std::string &v) { return storeCount(v, &c.sieveSegmentBytes); }},
    {"--sieve-blocks-per-cu", nullptr, true,
     [](Config &c, const std::string &v) { return storeCount(v, &c.sieveBlocksPerCU); }},
    {"--tune", nullptr, true,
     [](Config &c, const std::string &v) -> std::string {
         c.tuneFile = v;
         return "";
     }},
    {"--batch", nullptr, false,
     [](Config &c, const std::string &) -> std::string {
         c.mode = "batch";
         return "";
     }},
    {"--serve", nullptr, false,
     [](Config &c, const std::string &) -> std::string {
         c.mode = "serve";
         return "";
     }},
    {"--autotune", nullptr, false,
     [](Config &c, const std::string &) -> std::string {
         c.mode = "autotune";
         return "";
     }},
    {"--opt", nullptr, true,
     [](Config &c, const std::string &v) -> std::string {
         c.modeArgs.push_back(v);
         return "";
     }},
    // the device search is the only search here: --gpu-search is what
    // happens anyway, and there is no CPU search to fall back to
    {"--gpu-search", nullptr, false, [](Config &, const std::string &) -> std::string { return ""; }},
    {"--no-gpu", nullptr, false,
     [](Config &, const std::string &) -> std::string { return "this build searches on the GPU only"; }},
    // vendor selection of the reference's multi-vendor scheduler: one AMD
    // device here, so only the choices that keep it are accepted
    {"--devices", nullptr, true,
     [](Config &, const std::string &v) -> std::string {
         if (sameNoCase(v, "amd")) return "";
         return sameNoCase(v, "nvidia") ? "this build drives AMD (gfx950) devices only" : "not amd or nvidia";
     }},
    {nullptr, "FF_DISABLE_DEVICE", true,
     [](Config &, const std::string &v) -> std::string {
         if (sameNoCase(v, "nvidia")) return "";
         return sameNoCase(v, "amd") ? "disabling amd leaves no device" : "not amd or nvidia";
     }},
    {nullptr, "FF_PULL_WEIGHTS", true,
     [](Config &, const std::string &) -> std::string {
         return "pull weights balance several devices; this build drives one";
     }},
};

const Knob *findFlag(const std::string &name)
{
    for (const Knob &k : kKnobs)
        if (k.flag && name == k.flag) return &k;
    return nullptr;
}

int complain(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int complain(const char *fmt, ...)
{
    std::fputs("[ff_sieve] error: ", stderr);
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fputc('\n', stderr);
    return -1;
}

int fromEnvironment(Config *cfg)
{
    for (const Knob &k : kKnobs) {
        if (!k.env) continue;
        const char *v = std::getenv(k.env);
        if (!v || !*v) continue;
        const std::string why = k.apply(*cfg, v);
        if (!why.empty()) return complain("%s='%s': %s", k.env, v, why.c_str());
    }
    return 0;
}

int fromCommandLine(int argc, char **argv, Config *cfg, std::vector<std::string> *positionals)
{
    for (int i = 1; i < argc; ++i) {
        const std::string word = argv[i];
        if (word.compare(0, 2, "--") != 0) {
            positionals->push_back(word);
            continue;
        }
        // --flag=value or --flag value
        const size_t eq = word.find('=');
        const std::string name = word.substr(0, eq);
        const Knob *k = findFlag(name);
        if (!k) return complain("unknown option '%s'", name.c_str());
        std::string value;
        if (eq != std::string::npos) {
            if (!k->takesValue) return complain("%s takes no value", name.c_str());
            value = word.substr(eq + 1);
        } else if (k->takesValue) {
            if (i + 1 >= argc) return complain("%s needs a value", name.c_str());
            value = argv[++i];
        }
        const std::string why = k->apply(*cfg, value);
        if (!why.empty()) return complain("%s '%s': %s", name.c_str(), value.c_str(), why.c_str());
    }
    return 0;
}

// Checks that need the whole configuration.  Fractions of device memory
// stay within [0.10, 1.0]; an explicit budget or scratch is positive; a slab
// holds whole 16-value groups of the bitmap in bits, i.e. an even byte count
// (the reference's alignment rule); PRIMEMEM holds at least one map byte
// group.
int checkConfig(const Config &c)
{
    auto fractionOk = [](double f) { return f >= 0.10 && f <= 1.0; };
    if (!fractionOk(c.globalFraction))
        return complain("--vram-fraction %g: the fraction of free device memory must be in [0.10, 1.0]",
                        c.globalFraction);
    for (const auto &kv : c.deviceFractions)
        if (!fractionOk(kv.second))
            return complain("--device-vram-fraction %s=%g: must be in [0.10, 1.0]", kv.first.c_str(), kv.second);
    if (c.hasBudgetCap && !c.budgetCapBytes) return complain("--vram-budget 0: a budget must be positive");
    if (c.hasScratch && !c.scratchBytes) return complain("--scratch 0: scratch must be positive");
    if (!c.slabSizeBytes || (c.slabSizeBytes & 1))
        return complain("--slab-size %llu: a slab is a positive, even number of map bytes",
                        (unsigned long long)c.slabSizeBytes);
    if (c.primeMem < 16) return complain("--primemem %llu: below one map byte group", (unsigned long long)c.primeMem);
    return 0;
}

// The host tier's "auto" size: what the host's memory can hold with an
// eighth of it (at least 4 GiB) left for everything else.
uint64_t autoHostTier()
{
    const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGE_SIZE);
    if (pages <= 0 || page <= 0) return 0;
    const uint64_t ram = (uint64_t)pages * (uint64_t)page;
    const uint64_t keep = ram / 8 > (4ull << 30) ? ram / 8 : (4ull << 30);
    return ram > keep ? ram - keep : 0;
}

}  // namespace

bool readByteSize(const std::string &text, uint64_t *out)
{
    // number, then an optional unit: B; K, M or G with an optional i and an
    // optional B
    const char *s = text.c_str();
    if (!*s || !(std::isdigit((unsigned char)*s) || *s == '.')) return false;
    char *end = nullptr;
    const double count = std::strtod(s, &end);
    if (end == s || !(count >= 0.0)) return false;
    std::string unit(end);
    double scale = 1.0;
    if (!unit.empty()) {
        const char u = (char)std::toupper((unsigned char)unit[0]);
        const int shift = u == 'K' ? 10 : u == 'M' ? 20 : u == 'G' ? 30 : u == 'B' ? 0 : -1;
        if (shift < 0) return false;
        std::string rest = unit.substr(1);
        if (shift && !rest.empty() && (rest[0] == 'i' || rest[0] == 'I')) rest.erase(0, 1);
        if (shift && !rest.empty() && (rest[0] == 'b' || rest[0] == 'B')) rest.erase(0, 1);
        if (!rest.empty()) return false;
        scale = std::ldexp(1.0, shift);
    }
    const double bytes = count * scale;
    if (bytes >= std::ldexp(1.0, 64)) return false;
    uint64_t n = (uint64_t)std::nearbyint(bytes);
    if (n == 0 && bytes > 0.0) n = 1;  // a positive size never rounds to nothing
    *out = n;
    return true;
}

bool readFraction(const std::string &text, double *out)
{
    const char *s = text.c_str();
    if (!*s || !(std::isdigit((unsigned char)*s) || *s == '.')) return false;
    char *end = nullptr;
    const double v = std::strtod(s, &end);
    if (end == s || *end || !std::isfinite(v)) return false;
    *out = v;
    return true;
}

bool readFractionMap(const std::string &text, std::map<std::string, double> *out)
{
    std::map<std::string, double> m;
    size_t at = 0;
    while (at <= text.size()) {
        size_t stop = text.find(',', at);
        if (stop == std::string::npos) stop = text.size();
        const std::string item = text.substr(at, stop - at);
        const size_t eq = item.find('=');
        if (eq == std::string::npos || eq == 0) return false;
        double f = 0;
        if (!readFraction(item.substr(eq + 1), &f)) return false;
        m[item.substr(0, eq)] = f;
        at = stop + 1;
        if (stop == text.size()) break;
    }
    if (m.empty()) return false;
    *out = m;
    return true;
}

int loadTuneFile(Config *cfg, const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return complain("cannot read tuning file %s", path.c_str());
    std::string text;
    char buf[4096];
    for (size_t n; (n = std::fread(buf, 1, sizeof buf, f)) > 0;) text.append(buf, n);
    std::fclose(f);
    const struct {
        const char *key;
        uint32_t *dst;
    } keys[] = {{"search_grab", &cfg->searchGrab},
                {"search_eager_splits", &cfg->searchEager},
                {"search_blocks_per_cu", &cfg->searchBlocksPerCU},
                {"sieve_segment_bytes", &cfg->sieveSegmentBytes},
                {"sieve_blocks_per_cu", &cfg->sieveBlocksPerCU}};
    for (const auto &k : keys) {
        // the first occurrence of the key, then its number after any of
        // the separators of "key": n, "key" = n or key=n
        const size_t at = text.find(k.key);
        if (at == std::string::npos) continue;
        size_t p = at + std::strlen(k.key);
        while (p < text.size() && std::strchr("\": =", text[p])) ++p;
        size_t q = p;
        while (q < text.size() && std::isdigit((unsigned char)text[q])) ++q;
        uint64_t v = 0;
        if (q == p || !readUnsigned(text.substr(p, q - p), &v) || v > 0xFFFFFFFFull)
            return complain("tuning file %s: no number after %s", path.c_str(), k.key);
        if (*k.dst == 0) *k.dst = (uint32_t)v;  // a flag given on the command line wins
    }
    return 0;
}

int loadConfig(int argc, char **argv, Config *cfg, std::vector<std::string> *positionals)
{
    if (fromEnvironment(cfg) || fromCommandLine(argc, argv, cfg, positionals) || checkConfig(*cfg)) return -1;
    if (!cfg->tuneFile.empty() && loadTuneFile(cfg, cfg->tuneFile)) return -1;
    if (cfg->noHostTier) {
        cfg->hostTierAuto = false;
        cfg->hostTierCapBytes = 0;
    }
    if (cfg->hostTierAuto) {
        cfg->hostTierCapBytes = autoHostTier();
        cfg->hostTierAuto = false;
    }
    return 0;
}

}  // namespace ffg
