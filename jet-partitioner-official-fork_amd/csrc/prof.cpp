// Kernel-family timing (see prof.h).
#include "prof.h"

#include <vector>

#include "jet_types.h"

namespace jet {

Prof& prof() {
    static Prof* p = new Prof();
    return *p;
}

unsigned long long* prof_counters(int fam) { return prof_on(fam) ? prof().dev : nullptr; }

unsigned long long* prof_span(int fam) {
    if (!prof_on(fam)) return nullptr;
    Prof& P = prof();
    if (P.spans >= PROF_MAX_SPANS) return nullptr;
    return P.span + 2 * (P.spans++);
}

void prof_account(int fam, double host_bytes) {
    if (!prof_on(fam)) return;
    Prof& P = prof();
    P.host_bytes += host_bytes;
    P.launches++;
}

bool prof_enable(const std::string& name) {
    int fam = PF_NONE;
    if (name == "lp_select") fam = PF_LP_SELECT;
    else if (name == "conn_build") fam = PF_CONN_BUILD;
    else if (name == "project") fam = PF_PROJECT;
    else return false;
    Prof& P = prof();
    if (!P.dev) JET_HIP(hipMalloc(&P.dev, 8 * sizeof(unsigned long long)));
    if (!P.span) JET_HIP(hipMalloc(&P.span, 2 * PROF_MAX_SPANS * sizeof(unsigned long long)));
    JET_HIP(hipMemset(P.dev, 0, 8 * sizeof(unsigned long long)));
    // starts at all-ones (atomic min), ends at zero (atomic max)
    std::vector<unsigned long long> init(2 * PROF_MAX_SPANS);
    for (int64_t i = 0; i < PROF_MAX_SPANS; i++) {
        init[2 * i] = ~0ull;
        init[2 * i + 1] = 0;
    }
    JET_HIP(hipMemcpy(P.span, init.data(), init.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    JET_HIP(hipDeviceSynchronize());
    P.host_bytes = 0;
    P.launches = 0;
    P.spans = 0;
    P.family = fam;
    return true;
}

bool prof_read(int64_t& launches, double& ms, double& bytes) {
    Prof& P = prof();
    if (P.family == PF_NONE) return false;
    JET_HIP(hipDeviceSynchronize());
    int dev = 0, khz = 0;
    JET_HIP(hipGetDevice(&dev));
    JET_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    std::vector<unsigned long long> sp(2 * P.spans);
    if (P.spans > 0)
        JET_HIP(hipMemcpy(sp.data(), P.span, sp.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double ticks = 0;
    int64_t timed = 0;
    for (int64_t i = 0; i < P.spans; i++)
        if (sp[2 * i + 1] >= sp[2 * i] && sp[2 * i] != ~0ull) {
            ticks += (double)(sp[2 * i + 1] - sp[2 * i]);
            timed++;
        }
    unsigned long long c[8];
    JET_HIP(hipMemcpy(c, P.dev, sizeof(c), hipMemcpyDeviceToHost));
    bytes = P.host_bytes;
    if (P.family == PF_LP_SELECT) {
        // per cache miss: conn_off 8 + conn_size 4 + part 4 + gain write 4 +
        // cache write 4; per table slot 8; per candidate 4 (list write)
        bytes += 24.0 * (double)c[0] + 8.0 * (double)c[1] + 4.0 * (double)c[2];
    }
    // the span count bounds the launches whose time is known; scale the
    // byte total to the timed launches when spans ran out
    if (timed < P.launches && P.launches > 0) bytes *= (double)timed / (double)P.launches;
    launches = timed;
    ms = khz > 0 ? ticks / (double)khz : 0.0;
    return true;
}

}  // namespace jet
