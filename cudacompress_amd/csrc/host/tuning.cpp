// Tuning table of the kernel shapes (no reference counterpart).  The autotuner
// (cudacompress_amd/autotune.py) times candidate shapes on the device and writes the table; the
// library reads it once per process and every new instance starts from it.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>

#include "../common.hpp"
#include "instance.hpp"

namespace cca {

cc_tuning default_tuning() {
    cc_tuning t{};
    // resident design workgroups per class: LDS-bound occupancy of the four classes
    t.designGrid[0] = 4096;
    t.designGrid[1] = 2048;
    t.designGrid[2] = 1536;
    t.designGrid[3] = 768;
    t.fusedXY = 1;
    t.xySlicesMax = 0;
    t.codingBlockSize = 128;
    t.laneDesignK = 0;
    return t;
}

namespace {

std::string library_table_path() {
    if (const char* e = std::getenv("CUDACOMPRESS_AMD_TUNING")) return e;
    Dl_info info{};
    if (dladdr((void*)&default_tuning, &info) && info.dli_fname) {
        std::string p = info.dli_fname;  // .../cudacompress_amd/lib/libcudacompress_amd.so
        const size_t a = p.find_last_of('/');
        if (a != std::string::npos) {
            p = p.substr(0, a);  // .../lib
            const size_t b = p.find_last_of('/');
            p = (b == std::string::npos ? std::string(".") : p.substr(0, b)) + "/tuning/gfx950.txt";
            return p;
        }
    }
    return "tuning/gfx950.txt";
}

void check(const cc_tuning& t) {
    for (uint32_t g : t.designGrid) CCA_REQUIRE(g >= 1 && g <= (1u << 20), CC_ERR_INVALID_ARG, "design grid out of range");
    CCA_REQUIRE(t.fusedXY <= 1, CC_ERR_INVALID_ARG, "fusedXY must be 0 or 1");
    CCA_REQUIRE(t.xySlicesMax <= 4096, CC_ERR_INVALID_ARG, "xySlicesMax out of range");
    const uint32_t c = t.codingBlockSize;
    CCA_REQUIRE(c >= 32 && c <= 256 && (c & (c - 1)) == 0, CC_ERR_INVALID_ARG,
                "codingBlockSize must be a power of two in [32, 256]");
    const uint32_t k = t.laneDesignK;
    CCA_REQUIRE(k == 0 || k == 32 || k == 64 || k == 128 || k == 256, CC_ERR_INVALID_ARG,
                "laneDesignK must be 0, 32, 64, 128 or 256");
}

}  // namespace

cc_tuning load_tuning_file(const std::string& path) {
    cc_tuning t = default_tuning();
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return t;
    char line[512];
    int lineno = 0;
    while (std::fgets(line, sizeof line, f)) {
        lineno++;
        std::string s(line);
        const size_t h = s.find('#');
        if (h != std::string::npos) s.resize(h);
        std::istringstream in(s);
        std::string key;
        if (!(in >> key)) continue;
        bool ok = true;
        if (key == "design_grid") {
            for (uint32_t& g : t.designGrid) ok = ok && static_cast<bool>(in >> g);
        } else if (key == "fused_xy") {
            ok = static_cast<bool>(in >> t.fusedXY);
        } else if (key == "xy_slices_max") {
            ok = static_cast<bool>(in >> t.xySlicesMax);
        } else if (key == "coding_block_size") {
            ok = static_cast<bool>(in >> t.codingBlockSize);
        } else if (key == "lane_design_k") {
            ok = static_cast<bool>(in >> t.laneDesignK);
        }  // other keys (measurements, provenance) are ignored
        if (!ok) {
            std::fclose(f);
            throw Error{CC_ERR_INVALID_ARG, path + ":" + std::to_string(lineno) + ": malformed tuning entry"};
        }
    }
    std::fclose(f);
    check(t);
    return t;
}

// The library's table, read once (defaults when it is missing or malformed).
const cc_tuning& library_tuning() {
    static std::once_flag once;
    static cc_tuning t;
    std::call_once(once, [] {
        try {
            t = load_tuning_file(library_table_path());
        } catch (const Error&) {
            t = default_tuning();
        }
    });
    return t;
}

void validate_tuning(const cc_tuning& t) { check(t); }

}  // namespace cca
