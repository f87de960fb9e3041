// Kernel-family timing (see prof.h).
#include "prof.h"

#include "jet_types.h"

namespace jet {

Prof& prof() {
    static Prof* p = new Prof();
    return *p;
}

static hipEvent_t take_event() {
    Prof& P = prof();
    if (!P.pool.empty()) {
        hipEvent_t e = P.pool.back();
        P.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    JET_HIP(hipEventCreate(&e));
    return e;
}

unsigned long long* prof_counters(int fam) { return prof_on(fam) ? prof().dev : nullptr; }

std::pair<hipEvent_t, hipEvent_t> prof_events(int fam) {
    if (!prof_on(fam)) return {nullptr, nullptr};
    hipEvent_t a = take_event();
    hipEvent_t b = take_event();
    prof().pending.push_back({a, b});
    return {a, b};
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
    JET_HIP(hipMemset(P.dev, 0, 8 * sizeof(unsigned long long)));
    JET_HIP(hipDeviceSynchronize());
    for (auto& pr : P.pending) {
        P.pool.push_back(pr.first);
        if (pr.second) P.pool.push_back(pr.second);
    }
    P.pending.clear();
    P.host_bytes = 0;
    P.launches = 0;
    P.family = fam;
    return true;
}

bool prof_read(int64_t& launches, double& ms, double& bytes) {
    Prof& P = prof();
    if (P.family == PF_NONE) return false;
    JET_HIP(hipDeviceSynchronize());
    ms = 0;
    for (auto& pr : P.pending) {
        float t = 0;
        JET_HIP(hipEventElapsedTime(&t, pr.first, pr.second));
        ms += t;
    }
    unsigned long long c[8];
    JET_HIP(hipMemcpy(c, P.dev, sizeof(c), hipMemcpyDeviceToHost));
    bytes = P.host_bytes;
    if (P.family == PF_LP_SELECT) {
        // per cache miss: conn_off 8 + conn_size 4 + part 4 + gain write 4 +
        // cache write 4; per table slot 8; per candidate 4 (list write)
        bytes += 24.0 * (double)c[0] + 8.0 * (double)c[1] + 4.0 * (double)c[2];
    }
    launches = P.launches;
    return true;
}

}  // namespace jet
