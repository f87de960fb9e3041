// Process-wide tuning parameters (csrc/tuning.h), the per-architecture tuning
// table and their C-ABI accessors.
#include "tuning.h"

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <utility>
#include <vector>

#include "jet_capi.h"

namespace jet {

Tuning& tuning() {
    static Tuning t;
    return t;
}

namespace {
thread_local bool t_batch_thread = false;
}

void set_batch_thread(bool on) { t_batch_thread = on; }

int64_t wg_threshold() { return t_batch_thread ? tuning().batch_wg_max_n : tuning().wg_max_n; }

namespace {
struct Field {
    const char* name;
    int64_t Tuning::*member;
    int64_t lo, hi;
};
// every value in a range gives the same results; the ranges keep each kernel
// within the limits it is built for (LDS table loads of at most one half)
const Field kFields[] = {
    {"wg_max_n", &Tuning::wg_max_n, 0, 1 << 20},
    {"batch_wg_max_n", &Tuning::batch_wg_max_n, 0, 1 << 20},
    {"wg_big_deg", &Tuning::wg_big_deg, 0, 1 << 30},
    {"small_level_n", &Tuning::small_level_n, 0, 1 << 30},
    {"ab_lanes_small", &Tuning::ab_lanes_small, 1, 4},
    {"upd_lanes", &Tuning::upd_lanes, 1, 4},
    {"upd_big_slices", &Tuning::upd_big_slices, 1, 64},
    {"lp_one_pass_n", &Tuning::lp_one_pass_n, 0, int64_t(1) << 31},
    {"ct_wave_max", &Tuning::ct_wave_max, 1, 128},
    {"ct_block_max", &Tuning::ct_block_max, 1, 4096},
    {"redraw_lds_max", &Tuning::redraw_lds_max, 0, 8192},
    {"sel_max_blocks", &Tuning::sel_max_blocks, 64, 65536},
    {"sel_copy_wgs", &Tuning::sel_copy_wgs, 1, 65536},
};
std::mutex g_mu;
// fields set through jet_tuning_set: the table applied at context creation
// leaves them as the caller set them
uint32_t g_explicit = 0;

const Field* find_field(const char* name) {
    for (const auto& f : kFields)
        if (std::strcmp(f.name, name) == 0) return &f;
    return nullptr;
}

// Directory of this library (for the tuning tables installed beside it).
std::string library_dir() {
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&tuning_load), &info) == 0 || !info.dli_fname) return ".";
    std::string p = info.dli_fname;
    const size_t s = p.rfind('/');
    return s == std::string::npos ? "." : p.substr(0, s);
}
}  // namespace

static int load_table(const std::string& path, bool keep_explicit) {
    std::ifstream in(path);
    if (!in) return -1;
    std::vector<std::pair<const Field*, int64_t>> vals;
    std::string line;
    while (std::getline(in, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        std::string name;
        if (!(ss >> name)) continue;
        int64_t v = 0;
        const Field* f = find_field(name.c_str());
        if (!f || !(ss >> v) || v < f->lo || v > f->hi) return -1;
        vals.emplace_back(f, v);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& fv : vals) {
        const uint32_t bit = 1u << (fv.first - kFields);
        if (keep_explicit && (g_explicit & bit)) continue;
        tuning().*(fv.first->member) = fv.second;
    }
    return (int)vals.size();
}

int tuning_load(const std::string& path) { return load_table(path, false); }

void tuning_init_once(const std::string& arch) {
    static std::once_flag once;
    std::call_once(once, [&] {
        const char* e = std::getenv("JET_TUNING_FILE");
        const std::string path = e ? std::string(e) : library_dir() + "/tuning/" + arch + ".txt";
        (void)load_table(path, true);  // no table: the built-in defaults stay
    });
}

}  // namespace jet

extern "C" int jet_tuning_set(const char* name, int64_t value) {
    if (!name) return JET_EINVAL;
    const jet::Field* f = jet::find_field(name);
    if (!f || value < f->lo || value > f->hi) return JET_EINVAL;
    std::lock_guard<std::mutex> lk(jet::g_mu);
    jet::tuning().*(f->member) = value;
    jet::g_explicit |= 1u << (f - jet::kFields);
    return JET_OK;
}

extern "C" int jet_tuning_get(const char* name, int64_t* value) {
    if (!name || !value) return JET_EINVAL;
    const jet::Field* f = jet::find_field(name);
    if (!f) return JET_EINVAL;
    std::lock_guard<std::mutex> lk(jet::g_mu);
    *value = jet::tuning().*(f->member);
    return JET_OK;
}

extern "C" int jet_tuning_load(const char* path) {
    if (!path) return JET_EINVAL;
    return jet::tuning_load(path) < 0 ? JET_EINVAL : JET_OK;
}
