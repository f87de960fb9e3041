// Kernel-family timing and counting (see prof.h).
#include "prof.h"

#include <vector>

#include "jet_types.h"

namespace jet {

Prof& prof() {
    static Prof* p = new Prof();
    return *p;
}

unsigned long long* prof_counters(int fam) { return prof_on(fam) && prof().counting ? prof().dev : nullptr; }

unsigned long long* prof_span(int fam) {
    if (!prof_on(fam) || prof().counting) return nullptr;
    Prof& P = prof();
    if (P.spans >= PROF_MAX_SPANS) return nullptr;
    return P.span + 2 * PROF_SPAN_SLOTS * (P.spans++);
}

void prof_account(int fam, double host_bytes) {
    if (!prof_on(fam)) return;
    Prof& P = prof();
    P.host_bytes += host_bytes;
    P.launches++;
}

bool prof_enable(const std::string& spec) {
    std::string name = spec;
    bool counting = false;
    const std::string suffix = ":count";
    if (name.size() > suffix.size() && name.compare(name.size() - suffix.size(), suffix.size(), suffix) == 0) {
        counting = true;
        name.resize(name.size() - suffix.size());
    }
    if (name == "none") {
        prof().family = PF_NONE;
        return true;
    }
    int fam = PF_NONE;
    if (name == "lp_select") fam = PF_LP_SELECT;
    else return false;
    Prof& P = prof();
    const size_t words = 2 * PROF_SPAN_SLOTS * (size_t)PROF_MAX_SPANS;
    if (!P.dev) JET_HIP(hipMalloc(&P.dev, 8 * sizeof(unsigned long long)));
    if (!P.span) JET_HIP(hipMalloc(&P.span, words * sizeof(unsigned long long)));
    JET_HIP(hipMemset(P.dev, 0, 8 * sizeof(unsigned long long)));
    // starts at all-ones (atomic min), ends at zero (atomic max)
    std::vector<unsigned long long> init(words);
    for (size_t i = 0; i < words; i++) init[i] = (i / PROF_SPAN_SLOTS) % 2 == 0 ? ~0ull : 0ull;
    JET_HIP(hipMemcpy(P.span, init.data(), words * sizeof(unsigned long long), hipMemcpyHostToDevice));
    JET_HIP(hipDeviceSynchronize());
    P.host_bytes = 0;
    P.launches = 0;
    P.spans = 0;
    P.counting = counting;
    P.family = fam;
    return true;
}

bool prof_read(int64_t& launches, double& ms, double& bytes) {
    Prof& P = prof();
    if (P.family == PF_NONE) return false;
    JET_HIP(hipDeviceSynchronize());
    bytes = P.host_bytes;
    launches = P.launches;
    ms = 0;
    if (P.counting) {
        unsigned long long c[8];
        JET_HIP(hipMemcpy(c, P.dev, sizeof(c), hipMemcpyDeviceToHost));
        if (P.family == PF_LP_SELECT) {
            // per cache miss: conn_off 8 + conn_size 4 + part 4 + gain write 4 +
            // cache write 4; per table slot 8; per candidate: gain read 4,
            // pre-gain write 4, row offsets 16, list write 4
            bytes += 24.0 * (double)c[0] + 8.0 * (double)c[1] + 28.0 * (double)c[2];
        }
        return true;
    }
    int dev = 0, khz = 0;
    JET_HIP(hipGetDevice(&dev));
    JET_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    std::vector<unsigned long long> sp(2 * PROF_SPAN_SLOTS * (size_t)P.spans);
    if (P.spans > 0)
        JET_HIP(hipMemcpy(sp.data(), P.span, sp.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double ticks = 0;
    int64_t timed = 0;
    for (int64_t i = 0; i < P.spans; i++) {
        const unsigned long long* r = sp.data() + 2 * PROF_SPAN_SLOTS * i;
        unsigned long long s0 = ~0ull, e0 = 0;
        for (int q = 0; q < PROF_SPAN_SLOTS; q++) {
            s0 = r[q] < s0 ? r[q] : s0;
            e0 = r[PROF_SPAN_SLOTS + q] > e0 ? r[PROF_SPAN_SLOTS + q] : e0;
        }
        if (s0 != ~0ull && e0 >= s0) {
            ticks += (double)(e0 - s0);
            timed++;
        }
    }
    // launches beyond the span records are not timed: report the timed ones
    if (timed < P.launches && P.launches > 0) bytes *= (double)timed / (double)P.launches;
    launches = timed;
    ms = khz > 0 ? ticks / (double)khz : 0.0;
    return true;
}

}  // namespace jet
