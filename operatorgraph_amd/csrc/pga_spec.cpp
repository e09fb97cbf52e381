// pga_spec.cpp - kernels specialised to one program, compiled at run time.
//
// pga_spec_source writes the program as C++: one type per dispatch entry with
// the entry's fields as compile-time constants and a dispatcher from entry
// index to type (pga_spec.h runs the derivation over them). Programs of more
// than kSpecConstantEntriesMax entries are written as groups of entries that
// share their header words (operator, shape, counts, flags) instead: the
// compile time grows with the code, and a grammar of a few hundred entries
// has a few dozen operator forms. pga_spec_compile
// compiles that source with hipRTC for gfx950 with the flags of the static
// build (no FP contraction, correctly rounded division, no packed FP32), so
// the specialised kernel computes the interpreter's instance records bit for
// bit. Code objects are kept per process by a hash of their source, so
// contexts that load the same program compile it once.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "pga_spec_host.h"

// The device headers as text (pga.h, pga_device.h, pga_ops.h, pga_kernels.h,
// pga_level.h, pga_spec.h), written into the build tree by the build script.
#include "pga_rtc_headers.inc"

namespace {

uint64_t fnv1a(const std::string& s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : s) {
        h ^= ch;
        h *= 1099511628211ull;
    }
    return h;
}

std::mutex g_cache_mu;
std::unordered_map<uint64_t, std::vector<char>>& code_cache() {
    static std::unordered_map<uint64_t, std::vector<char>> cache;   // host memory only
    return cache;
}

std::string hex32(uint32_t v) {
    char b[16];
    std::snprintf(b, sizeof b, "0x%08Xu", v);
    return b;
}

// Programs up to this many entries get one constant type per entry.
constexpr size_t kSpecConstantEntriesMax = 32;

std::string fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return "__builtin_bit_cast(float, " + hex32(u) + ")";
}

}  // namespace

std::string pga_spec_source(const std::vector<pga::DevEntry>& entries, uint32_t count_geom, uint32_t use_seed,
                            uint32_t block_size) {
    const size_t n = entries.size();
    std::vector<uint64_t> key(n);
    for (size_t i = 0; i < n; ++i) {
        const pga::DevEntry& d = entries[i];
        const uint32_t w0 = (uint32_t)d.kind | ((uint32_t)d.shape << 8) | ((uint32_t)d.flags << 16) |
                            ((uint32_t)d.nparams << 24);
        const uint32_t w1 = (uint32_t)d.nsucc | ((uint32_t)d.inl << 8) | ((uint32_t)d.tslot << 16) |
                            ((uint32_t)d.ndirect << 24);
        key[i] = ((uint64_t)w1 << 32) | w0;
    }
    // entries per group key, in order of first appearance
    std::vector<uint64_t> groups;
    std::unordered_map<uint64_t, std::vector<size_t>> members;
    for (size_t i = 0; i < n; ++i) {
        auto& m = members[key[i]];
        if (m.empty()) groups.push_back(key[i]);
        m.push_back(i);
    }
    const bool grouped = n > kSpecConstantEntriesMax;

    std::string s;
    s.reserve(4096 + n * 400);
    s += "// derivation kernel specialised to one program (written by pga_spec_source)\n";
    s += "typedef __hip_internal::int8_t int8_t;\ntypedef __hip_internal::uint8_t uint8_t;\n"
         "typedef __hip_internal::int16_t int16_t;\ntypedef __hip_internal::uint16_t uint16_t;\n"
         "typedef __hip_internal::int32_t int32_t;\ntypedef __hip_internal::uint32_t uint32_t;\n"
         "typedef __hip_internal::int64_t int64_t;\ntypedef __hip_internal::uint64_t uint64_t;\n";
    s += "#include \"pga_spec.h\"\nnamespace pga {\nnamespace gen {\n";
    // one constant type per entry (per entry alone in its group, when grouped)
    for (size_t i = 0; i < n; ++i) {
        if (grouped && members[key[i]].size() > 1) continue;
        const pga::DevEntry& d = entries[i];
        const std::string nm = "E" + std::to_string(i);
        s += "struct " + nm + " : SpecEntry<" + nm + "> {\n";
        s += "  static constexpr uint32_t idx = " + std::to_string(i) + "u, w0 = " + hex32((uint32_t)key[i]) +
             ", w1 = " + hex32((uint32_t)(key[i] >> 32)) + ", pbase = " + std::to_string(d.pbase) +
             "u, sbase = " + std::to_string(d.sbase) + "u, aux0 = " + hex32(d.aux[0]) + ", aux1 = " +
             hex32(d.aux[1]) + ";\n";
        s += "  static constexpr float pv0 = " + fbits(d.pv[0]) + ", pv1 = " + fbits(d.pv[1]) +
             ", pv2 = " + fbits(d.pv[2]) + ";\n";
        s += "  static constexpr int32_t sv0 = " + std::to_string(d.sv[0]) + ", sv1 = " + std::to_string(d.sv[1]) +
             ", sv2 = " + std::to_string(d.sv[2]) + ";\n};\n";
    }
    s += "struct Prog {\n";
    bool polygons = false;   // a dynamic polygon or prism procedure: 6, 7, 14, 15
    for (const pga::DevEntry& d : entries)
        polygons |= d.shape == PGA_SHAPE_DYNAMIC_CONVEX_POLYGON || d.shape == PGA_SHAPE_DYNAMIC_POLYGON ||
                    d.shape == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM || d.shape == PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
    s += "  static constexpr uint32_t kCountGeom = " + std::to_string(count_geom) +
         "u, kUseSeed = " + std::to_string(use_seed) + "u, kEntries = " + std::to_string(n) +
         "u, kPolygons = " + (polygons ? "1u" : "0u") + ";\n";
    s += std::string("  static constexpr bool kGrouped = ") + (grouped ? "true" : "false") + ";\n";
    s += "  template <class F> __device__ __forceinline__ static void dispatch(const Ctx& c, uint64_t u, uint32_t i, "
         "F& f) {\n";
    s += "    switch (u) {\n";
    if (!grouped) {
        for (size_t i = 0; i < n; ++i)
            s += "    case " + std::to_string(i) + "ull: f(E" + std::to_string(i) + "{}); return;\n";
    } else {
        for (uint64_t g : groups) {
            const auto& m = members[g];
            char k[32];
            std::snprintf(k, sizeof k, "0x%016llXull", (unsigned long long)g);
            if (m.size() == 1)
                s += std::string("    case ") + k + ": f(E" + std::to_string(m[0]) + "{}); return;\n";
            else
                s += std::string("    case ") + k + ": f(SpecGroupEntry<" + hex32((uint32_t)g) + ", " +
                     hex32((uint32_t)(g >> 32)) + ">::at(c, i)); return;\n";
        }
    }
    s += "    default: f(SpecNoEntry{}); return;\n    }\n  }\n};\n";
    s += "}  // namespace gen\n}  // namespace pga\n";
    s += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(block_size) +
         ") pga_spec_mega(pga::KArgs a) {\n  pga::spec_mega<pga::gen::Prog, " + std::to_string(block_size) +
         ">(a);\n}\n";
    return s;
}

bool pga_spec_compile(const std::string& source, std::vector<char>& code, std::string& log, bool* cached) {
    const uint64_t key = fnv1a(source);
    if (cached) *cached = false;
    if (const char* dir = std::getenv("PGA_SPEC_DUMP")) {   // diagnostics: keep the generated source
        char name[64];
        std::snprintf(name, sizeof name, "/pga_spec_%016llx.hip", (unsigned long long)key);
        if (FILE* f = std::fopen((std::string(dir) + name).c_str(), "w")) {
            std::fwrite(source.data(), 1, source.size(), f);
            std::fclose(f);
        }
    }
    {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        auto it = code_cache().find(key);
        if (it != code_cache().end()) {
            code = it->second;
            if (cached) *cached = true;
            return true;
        }
    }
    const size_t nh = sizeof(kRtcHeaderNames) / sizeof(kRtcHeaderNames[0]);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source.c_str(), "pga_spec.hip", (int)nh, kRtcHeaderTexts, kRtcHeaderNames) !=
        HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                          "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    log.assign(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    bool ok = r == HIPRTC_SUCCESS;
    if (ok) {
        size_t cs = 0;
        ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
        if (ok) {
            code.resize(cs);
            ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
        }
    }
    hiprtcDestroyProgram(&prog);
    if (ok) {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        code_cache()[key] = code;
    }
    return ok;
}
