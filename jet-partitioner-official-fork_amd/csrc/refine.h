// Jet refinement on the GPU (jet_refiner.hpp and uncoarsen.hpp of the reference).
#pragma once
#include <array>
#include <vector>

#include "arena.h"
#include "contract.h"
#include "jet_types.h"
#include "refine_wg.h"

namespace jet {

// Per-level refinement record (experiment_data::CoarseLevel, experiment_data.hpp:102-123).
struct LevelRecord {
    int64_t cut2 = 0;       // best_state.cut after the level (2 x cut)
    int64_t total_imb = 0;  // best_state.total_imb after the level
    int32_t n = 0;
    int64_t m = 0;
    int32_t iterations = 0;
    int32_t lp_iterations = 0;
    double seconds = 0;
};

struct RefineParams {
    int32_t k = 2;
    double imb = 1.03;
    double tol = 0.999;
    bool ultra = false;
};

struct MoveRec;
struct BestCopy;

// Holds the refiner's scratch (sized by the finest level, jet_refiner.hpp:134-181)
// and the refine_data persisting across levels (jet_refiner.hpp:100-106).
class Refiner {
  public:
    Refiner(Arena& a, hipStream_t s, vid_t finest_n, int64_t finest_conn_cap, const RefineParams& p);
    ~Refiner() = default;
    Refiner(const Refiner&) = delete;
    Refiner& operator=(const Refiner&) = delete;

    // jet_refiner.hpp:1050-1142 on one level; best_part (device, L.n) in/out.
    LevelRecord refine(const Level& L, part_t* best_part);

    // One move-selection step on fresh state (empty cache, no locks) for
    // parity tests: kind 0 = jet_lp, 1 = rebalance_weak, 2 = rebalance_strong.
    // Returns moves (sorted by id) and their destinations on the host.
    void single_step(const Level& L, const part_t* part, int kind, double filter_ratio, std::vector<int32_t>& moves,
                     std::vector<int32_t>& dest);

    // Communication volume and cut weight per part of `part` (part_stat.hpp).
    void quality(const Level& L, const part_t* part, int64_t& comm_size, std::vector<int64_t>& cut_per_part);
    // cut weight between every ordered pair of parts, k * k entries (device)
    void heatmap(const Level& L, const part_t* part, int64_t* heat);

    // Connectivity tables for `part` (init_conn_data), copied to the host.
    void conn_tables(const Level& L, const part_t* part, std::vector<int64_t>& off, std::vector<int32_t>& slot_part,
                     std::vector<int32_t>& slot_val, std::vector<int32_t>& size);

    // When set, refine() appends one record per iteration: step kind (0 lp,
    // 1 weak, 2 strong), -1 (move count, not read back), cut2 and imbalance
    // after the step, best-state decision (0 none, 1 balance, 2 cut) and the
    // counter after it. A traced level runs on the multi-kernel path.
    std::vector<std::array<int64_t, 6>>* trace = nullptr;

    // Whether the last refine() replaced the best partition (and so the
    // refine_data carried to the next level).
    bool changed_last() const { return changed_; }
    int64_t cut2() const { return cut_; }
    int64_t total_size() const { return total_size_; }
    int64_t total_imb() const { return total_imb_; }
    int32_t opt() const { return opt_; }
    int32_t size_max() const { return size_max_; }
    const int64_t* best_sizes_dev() const { return best_sizes_; }

  private:
    // jet_refine on one level by the one-workgroup kernel (small levels)
    LevelRecord refine_wg(const Level& L, part_t* best_part, const std::vector<double>& temps);
    void init_conn(const Level& L, const part_t* part, const int32_t* mark);
    void build_tiers(const Level& L);
    void jet_lp(const Level& L, const part_t* part, double filter_ratio);
    void rebalance(const Level& L, const part_t* part, const int64_t* sizes, bool strong);
    // Bucket layout of one rebalancing step (jet_refiner.hpp:417-428, 551-560).
    struct RebalanceShape {
        int32_t sections = 0;
        int64_t width = 0, nbk = 0, max_dest = 0;
        int key_bits = 1;
    };
    RebalanceShape rebalance_shape(vid_t n, bool strong) const;
    void rank_flagged(const Level& L, const part_t* part, const int64_t* sizes, int64_t width, bool strong);
    void rank_radix(const Level& L, const part_t* part, const int64_t* sizes, int64_t width, int key_bits,
                    bool strong);
    // cookie_done: the ranking workgroup already ran the cookie cutter
    void strong_tail(const Level& L, const part_t* part, const int64_t* sizes, const RebalanceShape& r,
                     bool cookie_done);
    void rank_deferred(const Level& L, const part_t* part, const int64_t* sizes, bool strong);
    // Returns true when the step's ranking was deferred (nothing applied).
    void wait_published(volatile int64_t* res, int64_t seq);
    // rebuilds the rows an update_large step marked, if any
    void flush_rebuild(const Level& L);
    // next_strong >= 0: the second pass also prepares the start of a next
    // rebalancing step of that kind (strong 1, weak 0), so it needs no reset
    bool perform_moves(const Level& L, part_t* part, int64_t* sizes, int64_t& cut_change, int64_t& max_size,
                       int next_strong = -1);
    // Two counter sets alternate between iterations: the first kernel of an
    // iteration clears the other set, so no fill launch is needed.
    int32_t* cnt() const { return counters_ + 8 * par_; }
    int32_t* cnt_next() const { return counters_ + 8 * (1 - par_); }

    Arena& a_;
    hipStream_t s_;
    RefineParams p_;
    vid_t nmax_;
    // connectivity tables
    eoff_t* conn_off_;
    int2* conn_;
    int32_t* conn_size_;
    part_t* dest_cache_;
    int32_t* lock_;
    int32_t* gain_p_;
    // iteration scratch
    int32_t* pregain_;
    part_t* dest_;
    int32_t* bid_;
    vid_t* cand_;
    vid_t* moves_;
    MoveRec* mrec_;  // per listed move: row, vertex and new part (first pass to second)
    int32_t* swap_bit_;  // update_large: rows marked for rebuilding (swap_bit, jet_refiner.hpp:747-772)
    uint32_t* sval_;
    uint32_t* vtmp_;
    int64_t* prefix_;
    int64_t* score_;
    part_t* cur_part_;
    vid_t* tier_mid_;
    vid_t* tier_big_;
    // rows of degree > BIG_DEG ("hubs": tier_big_ and pull_mid_) take their
    // update_small changes through k_update_hubs: flags per vertex for hub
    // and hub with a moved neighbour
    uint8_t* hub_;
    uint8_t* hub_dirty_;
    // stamp_[v] == gen_: v is in the current step's move list (the move-list
    // producers write it; every selection step takes a new generation)
    int32_t* stamp_;
    int32_t gen_ = 0;
    // the last update pass prepared a rebalancing step of kind prep_strong_
    bool prep_ready_ = false, prep_strong_ = false;
    // a best-state copy (of cur_part_ and curr_sizes_ into best_out_ and
    // best_sizes_) is due: the next selection kernel or the level's end makes it
    bool best_due_ = false;
    part_t* best_out_ = nullptr;
    BestCopy take_best_copy();
    vid_t* pull_mid_;  // mid-tier rows of degree > BIG_DEG
    int32_t ntier_mid_ = 0, ntier_big_ = 0, npull_mid_ = 0;
    bool has_big_ = false;
    // small device state; counter set layout: [0] small candidates, [1] big
    // candidates, [2] small moves, [3] big moves, [4] re-draws, [5] ranked
    // bucket members, [6] update_large chosen
    int32_t* counters_;
    int par_ = 0;
    int64_t* acc_;  // [3] strong capacity / re-draw counter
    int64_t* curr_sizes_;
    int64_t* best_sizes_;
    int64_t* evict_adjust_;
    int64_t* evict_start_;
    part_t* undersized_;
    int32_t* misc_;  // [0] undersized count, [1] error bits, [2] ranking deferred
    vid_t* redraw_;    // strong re-draws (n)
    int32_t* rd_mark_; // marks of a long re-draw list (n, zero between steps)
    int64_t* buckets_;
    int64_t* boff_;        // the scanned bucket array (by the classification kernel when it is small)
    const int64_t* off_ = nullptr;  // the current rebalance step's scanned buckets (boff_)
    int32_t small_max_ = 0;         // rows of degree <= this take the thread-per-row table build
    bool rebuild_pending_ = false;  // update_large marked rows still to rebuild
    Mailbox* mailbox_ = nullptr;    // the creating thread's mailbox (results and sequence words)
    int64_t* pub_ = nullptr;      // the cut change: PUB_GROUPS partials, 128 bytes apart
    int64_t* szc_ = nullptr;      // the first update pass's part-size deltas: up to 16 copies of k
    int64_t* bcopies_ = nullptr;  // ncopies_ - 1 more bucket arrays, kept zero between uses
    int32_t ncopies_ = 1;
    int64_t nbuckets_ = 0;
    WgState* wg_state_ = nullptr;  // the one-workgroup kernel's controller state
    vid_t* wg_list_ = nullptr;     // its ranked-member list (n)
    int32_t* wg_runb_ = nullptr;   // its per-bucket running weights (WG_MAX_NBK)
    int64_t* res_host_ = nullptr;  // thread mailbox: (cut change, largest part) of the last iteration
    int64_t* res_dev_ = nullptr;
    // refine_data mirror on the host
    bool init_ = false;
    bool changed_ = false;
    int64_t cut_ = 0, total_imb_ = 0, total_size_ = 0;
    int32_t opt_ = 0, size_max_ = 0;
};

// uncoarsen.hpp:82-86.
void project(vid_t n_fine, const vid_t* map, const part_t* coarse, part_t* fine, hipStream_t s);

// part_stat.hpp:77-108 (2 x cut), :146-153.
int64_t total_cut2(const DGraph& g, const part_t* part, Arena& a, hipStream_t s);
// Adds the directed cut (g.ew NULL: unit weights) into *out (device).
void launch_cut2(const DGraph& g, const part_t* part, unsigned long long* out, hipStream_t s);
void part_sizes(const DGraph& g, const part_t* part, int32_t k, int64_t* sizes, hipStream_t s);

}  // namespace jet
