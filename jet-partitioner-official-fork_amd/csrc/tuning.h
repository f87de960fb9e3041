// Launch-shape and path parameters that do not change results (every setting
// gives the same partition): process-wide, read from the device's tuning
// table at context creation, settable through jet_tuning_set (the autotuner
// and the tests use it).
#pragma once
#include <cstdint>
#include <string>

namespace jet {

struct Tuning {
    // levels of up to this many vertices are refined by the one-workgroup
    // kernel (0: never)
    int64_t wg_max_n = 0;
    // the same threshold for jobs of the batched and served layer on contexts
    // of at least BATCH_WG_MIN_WORKERS workers, whose many concurrent small
    // jobs are bound by kernel launches more than by kernel time (their
    // one-workgroup levels are launched together, one launch per round)
    int64_t batch_wg_max_n = 0;
    // in that kernel, candidates and moves of higher degree are taken by a
    // whole wave instead of one thread
    int64_t wg_big_deg = 32;
    // "small" levels: every connectivity row is built by a wave, the
    // afterburner takes ab_lanes_small times its lanes per candidate, and the
    // rebalancers' select kernels run 128-thread workgroups
    int64_t small_level_n = 32768;
    int64_t ab_lanes_small = 2;
    // lanes per move in the update passes: upd_lanes times the level's lane
    // group (from its average degree), at most a wave
    int64_t upd_lanes = 2;
    // moves of degree above 64 are spread over this many waves each
    int64_t upd_big_slices = 16;
    // levels of up to this many vertices run LP select with one pass per wave
    int64_t lp_one_pass_n = 524288;
    // contraction: coarse rows of up to ct_wave_max fine entries take the
    // one-wave LDS table (256 slots), up to ct_block_max the one-workgroup
    // table (8192 slots), longer rows the radix spill path
    int64_t ct_wave_max = 128;
    int64_t ct_block_max = 4096;
    // strong rebalancing: up to this many re-drawn vertices are put in id
    // order in LDS, longer lists in global memory
    int64_t redraw_lds_max = 8192;
    // rebalancers' select kernels: at most this many workgroups (each a
    // contiguous range of vertices)
    int64_t sel_max_blocks = 1024;
    // and flush their bucket maps into one copy of the bucket array per this
    // many workgroups (at most 16 copies)
    int64_t sel_copy_wgs = 256;
};

Tuning& tuning();
// Contexts with fewer workers gather too few levels per launch for the
// one-workgroup kernel to pay: their jobs refine as single calls do.
constexpr int BATCH_WG_MIN_WORKERS = 8;
// Marks the calling thread as a worker of the batched and served layer.
void set_batch_thread(bool on);
// The one-workgroup refinement threshold for the calling thread: wg_max_n,
// or batch_wg_max_n on batch worker threads.
int64_t wg_threshold();

// Applies a tuning table (lines "name value"; '#' starts a comment). Returns
// the number of settings applied, or -1 when the file cannot be read or holds
// an unknown name or a value out of range (nothing is applied then).
int tuning_load(const std::string& path);

// Loads the table for the device architecture `arch` (e.g. "gfx950") once per
// process: $JET_TUNING_FILE when set, else tuning/<arch>.txt beside the
// library. Parameters already set through jet_tuning_set keep their values.
void tuning_init_once(const std::string& arch);

}  // namespace jet
