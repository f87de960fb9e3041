// Whole-level refinement in one workgroup, for small levels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "jet_types.h"

namespace jet {

// Controller state of jet_refine (jet_refiner.hpp:1079-1134) kept on the
// device between launches of the one-workgroup kernel. The kernel runs
// iterations until the level ends (status WG_DONE) or until a step chose
// update_large (status WG_REBUILD): the host then rebuilds the marked rows'
// tables with the multi-workgroup build kernels and launches it again, and it
// continues with the next iteration.
struct WgState {
    int64_t curr_cut, curr_imb, best_cut, best_imb;
    int32_t temp_idx, count, balance_counter, iterations, lp_iterations;
    int32_t changed, status, err;
};
enum : int32_t { WG_DONE = 0, WG_REBUILD = 1, WG_STALLED = 2 };

constexpr int WG_MAX_TEMPS = 20;

struct WgParams {
    DGraph g;
    int32_t k, sections;
    int32_t big_deg;  // items of higher degree are taken by a whole wave
    int32_t lanes;    // lanes per smaller item: a power of two up to a wave
    int32_t nbk;  // minibuckets scanned (the weak layout, the larger of the two)
    int64_t opt, size_max, imb_max, max_dest_weak, max_dest_strong;
    double tol;
    int32_t ntemps;
    double temps[WG_MAX_TEMPS];
    part_t* part;
    part_t* best_part;
    const eoff_t* conn_off;
    int2* conn;
    int32_t* conn_size;
    part_t* dest_cache;
    int32_t* lock;
    int32_t* gain_p;
    int32_t* pregain;
    part_t* dest;
    int32_t* bid;
    vid_t* cand;    // n: candidates (small from the front, large from the back)
    vid_t* moves;   // n: moves, the same layout
    vid_t* rlist;   // n: ranked bucket members in id order, then the re-draws
    int64_t* score; // n
    int32_t* mark;  // n: rows whose tables are to be rebuilt
    int32_t* runb;  // nbk: per-bucket running weight across ranking chunks
    int64_t* curr_sizes;
    int64_t* best_sizes;
    WgState* st;
};

// Levels the kernel takes: k and the bucket array bounded by its LDS.
constexpr int32_t WG_MAX_K = 256;
constexpr int32_t WG_MAX_NBK = 16384;

// Whether a level of n vertices with k parts fits the kernel.
bool wg_fits(int64_t n, int32_t k, int64_t nbk, int64_t max_n);
// Fresh controller state for a level: current and best cut / imbalance.
void wg_init_state(WgState* st, int64_t cut, int64_t imb, hipStream_t s);
// One launch of the kernel on stream s.
void launch_refine_wg(const WgParams& p, hipStream_t s);
// Gathers the one-workgroup level runs of concurrent jobs into graph-tagged
// launches (the batched and served layer has one per context). run() first
// waits for the job's stream (the level's tables and state are built there),
// and returns the level's final state once its workgroup has finished.
class WgCollector {
  public:
    virtual ~WgCollector() = default;
    virtual void run(const WgParams& p, hipStream_t job_stream, WgState& out) = 0;
};
// The collector of the calling thread's level runs (none: each run is its own
// launch on the job's stream).
void set_wg_collector(WgCollector* c);
WgCollector* wg_collector();

// One graph-tagged launch: workgroup i runs the level of dev_ps[i] (a device
// copy of host_ps[i]) and writes its final state to dev_states[i].
void launch_refine_wg_multi(const WgParams* host_ps, const WgParams* dev_ps, WgState* dev_states, int count,
                            hipStream_t s);

}  // namespace jet
