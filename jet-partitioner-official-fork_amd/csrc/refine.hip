// Jet refinement for gfx950 (reference: src/jet_refiner.hpp, src/part_stat.hpp,
// src/uncoarsen.hpp).
//
// Connectivity tables keep the reference's mechanics (capacity min(deg, k),
// direct mapping when the table size is k, linear probing from p % size
// otherwise, shrink to used + max(used/4, 3), HASH_RECLAIM slots, dense
// overflow append) so their contents, sizes and zero entries evolve exactly
// as in the reference. Their slot order is not kept: concurrent updates here
// place new parts in whichever free slot they reach first. Every decision
// that reads a table is therefore made independent of slot order: among parts
// with equal connectivity the lowest part id wins. The reference keeps the
// first maximum in slot order (jet_refiner.hpp:226-233, :603-611). For
// direct-mapped rows (table size k) the two rules agree. For hashed rows
// (size < k) they differ whenever the tied parts' slots are not in id order,
// e.g. parts 5 and 2 in a 4-slot table sit in slots 1 and 2 and the reference
// takes 5; such rows are built serially in edge order by the reference when
// the average degree is below 8, so there its choice is fixed and the lowest-id
// choice is not one the reference can make. Rebalancer minibucket positions
// follow increasing vertex id (a stable sort by bucket id), and the strong
// rebalancer's destination re-draws run in increasing vertex id.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <sched.h>

#include "prims.h"
#include "prof.h"
#include "conn_table.h"
#include "refine.h"
#include "tuning.h"
#include "wave.h"

namespace jet {

constexpr int32_t SMALL_DEG = 16;       // rows built by one lane
constexpr int32_t WAVE_TABLE_MAX = 1024; // rows with capacity <= this built by one wave
constexpr int32_t BIG_DEG = 64;         // candidates/moves above this degree use one wave
constexpr int32_t MAX_K = 8192;
constexpr int32_t LDS_K = 2048;         // part-size deltas aggregated in LDS up to this k
constexpr int32_t HEAVY_DEG_REFINE = 2048;  // rows above this degree are built by a workgroup

// ---------------------------------------------------------------- helpers
__device__ inline vid_t list_at(const vid_t* list, int32_t nsmall, vid_t nmax, int64_t x) {
    return x < nsmall ? list[x] : list[nmax - 1 - (x - nsmall)];
}

// ---------------------------------------------------------------- statistics
__global__ void k_cut2(DGraph g, const part_t* part, unsigned long long* out) {
    __shared__ int64_t red[256 / WAVE];
    __shared__ EdgeWalkLds walk;
    int64_t c = 0;
    block_edges<256>(g.rowptr, g.n, g.m, walk, [&](vid_t r, eoff_t j) {
        if (part[g.col[j]] != part[r]) c += g.ew ? g.ew[j] : 1;
    });
    c = block_reduce_sum<int64_t, 256>(c, red);
    if (threadIdx.x == 0 && c) atomicAdd(out, (unsigned long long)c);
}

void launch_cut2(const DGraph& g, const part_t* part, unsigned long long* out, hipStream_t s) {
    if (g.n > 0 && g.m > 0) {
        k_cut2<<<grid_for(g.m, EB_TILE, 8192), 256, 0, s>>>(g, part, out);
        JET_LAUNCH_CHECK();
    }
}

int64_t total_cut2(const DGraph& g, const part_t* part, Arena& a, hipStream_t s) {
    ArenaScope sc(a);
    unsigned long long* out = a.alloc<unsigned long long>(1);
    fill_i64(reinterpret_cast<int64_t*>(out), 0, 1, s);
    launch_cut2(g, part, out, s);
    return (int64_t)read_scalar(out, s);
}

__global__ void k_part_sizes(DGraph g, const part_t* part, int32_t k, unsigned long long* sizes) {
    __shared__ unsigned long long ls[LDS_K];
    bool lds = k <= LDS_K;
    if (lds) {
        for (int t = threadIdx.x; t < k; t += blockDim.x) ls[t] = 0;
        __syncthreads();
    }
    FOR_THREAD(i, g.n) {
        if (lds) atomicAdd(&ls[part[i]], (unsigned long long)g.vw[i]);
        else atomicAdd(&sizes[part[i]], (unsigned long long)g.vw[i]);
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < k; t += blockDim.x)
            if (ls[t]) atomicAdd(&sizes[t], ls[t]);
    }
}

void part_sizes(const DGraph& g, const part_t* part, int32_t k, int64_t* sizes, hipStream_t s) {
    fill_i64(sizes, 0, k, s);
    if (g.n > 0) {
        k_part_sizes<<<grid_for(g.n, 256, 2048), 256, 0, s>>>(g, part, k, reinterpret_cast<unsigned long long*>(sizes));
        JET_LAUNCH_CHECK();
    }
}

__global__ void k_project(vid_t n, const vid_t* map, const part_t* coarse, part_t* fine) {
    FOR_THREAD(i, n) fine[i] = coarse[map[i]];
}

// A best-state copy due from the previous step, made by the next step's
// selection kernel as it reads every vertex's part (part_out null: none due).
struct BestCopy {
    part_t* part_out;
    int64_t* sizes_out;
    const int64_t* sizes;
    int32_t k;
    __device__ void sizes_block0() const {
        if (part_out && blockIdx.x == 0)
            for (int32_t p = threadIdx.x; p < k; p += blockDim.x) sizes_out[p] = sizes[p];
    }
};

// The controller's best state (jet_refiner.hpp:1117-1132): the partition
// and the part sizes copied in one launch.
__global__ void k_copy_best(part_t* best_part, const part_t* part, vid_t n, int64_t* best_sizes,
                            const int64_t* sizes, int32_t k) {
    FOR_THREAD(i, n) best_part[i] = part[i];
    if (blockIdx.x == 0)
        for (int32_t p = threadIdx.x; p < k; p += blockDim.x) best_sizes[p] = sizes[p];
}

void project(vid_t n_fine, const vid_t* map, const part_t* coarse, part_t* fine, hipStream_t s) {
    if (n_fine <= 0) return;
    k_project<<<grid_for(n_fine, 256, 16384), 256, 0, s>>>(n_fine, map, coarse, fine);
    JET_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- table build
// jet_refiner.hpp:958-1048 and :648-739: capacity min(deg, k); direct slot p
// when the capacity is k, else linear probing from p % cap; shrink to
// used + max(used/4, 3) when that is smaller. The whole capacity is written
// (layout followed by empty slots). `mark` (optional) restricts the build to
// marked rows, clears their mark and their destination cache (update_large,
// :743-777).

__global__ void k_conn_caps(DGraph g, int32_t k, int32_t* caps) {
    FOR_THREAD(i, g.n) {
        int64_t d = g.rowptr[i + 1] - g.rowptr[i];
        caps[i] = (int32_t)(d > k ? k : d);
    }
}

// Table-build tiers (mid: one wave per row, big: one workgroup per row) and
// the rows whose update_small changes are pulled (degree > BIG_DEG): big-tier
// rows and pull_mid rows; hub[] flags both kinds.
__global__ void __launch_bounds__(256) k_conn_tiers(DGraph g, int32_t k, vid_t* mid, vid_t* big, vid_t* pull_mid,
                                                    int32_t* cnt, uint8_t* hub, int32_t small_max) {
    __shared__ int32_t sh[256 / WAVE + 1];
    FOR_BLOCK_UNIFORM(i, g.n) {
        const bool in = i < g.n;
        int64_t d = in ? g.rowptr[i + 1] - g.rowptr[i] : 0;
        int64_t cap = d > k ? k : d;
        bool is_mid = d > small_max && d <= HEAVY_DEG_REFINE && cap <= WAVE_TABLE_MAX;
        bool is_big = d > SMALL_DEG && !is_mid;
        bool is_pmid = is_mid && d > BIG_DEG;
        if (in) hub[i] = (is_big || is_pmid) ? 1 : 0;
        int32_t a = block_append<256>(&cnt[0], is_mid, sh);
        int32_t b = block_append<256>(&cnt[1], is_big, sh);
        int32_t c = block_append<256>(&cnt[3], is_pmid, sh);
        if (a >= 0) mid[a] = (vid_t)i;
        if (b >= 0) big[b] = (vid_t)i;
        if (c >= 0) pull_mid[c] = (vid_t)i;
    }
}

// Counting runs of the profiler only: adds the degrees of the items of a
// list (counters[cs] from the front, counters[cb] from the back of its nmax
// slots) to out[0]; returns at once when *deferred is set.
__global__ void __launch_bounds__(256)
    k_prof_degsum(DGraph g, const vid_t* list, const int32_t* counters, int cs, int cb, vid_t nmax,
                  const int32_t* deferred, unsigned long long* out) {
    __shared__ int64_t red[256 / WAVE];
    if (deferred && *deferred) return;
    const int64_t ns = counters[cs], nb = counters[cb];
    int64_t sum = 0;
    FOR_THREAD(x, ns + nb) {
        const vid_t v = x < ns ? list[x] : list[nmax - 1 - (x - ns)];
        sum += g.rowptr[v + 1] - g.rowptr[v];
    }
    sum = block_reduce_sum<int64_t, 256>(sum, red);
    if (threadIdx.x == 0 && sum) atomicAdd(out, (unsigned long long)sum);
}

constexpr int SB_BLOCK = 256;

__global__ void __launch_bounds__(SB_BLOCK)
    k_conn_build_small(DGraph g, const part_t* part, int32_t k, const eoff_t* conn_off, int2* conn,
                       int32_t* conn_size, int32_t* mark, part_t* dest_cache, const int32_t* gate, int32_t small_max) {
    if (gate && gate[0] == 0) return;
    __shared__ int2 tab[SMALL_DEG][SB_BLOCK];
    const int t = threadIdx.x;
    FOR_THREAD(i, g.n) {
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        int32_t deg = (int32_t)(re - rs);
        if (deg > small_max) continue;
        if (mark) {
            if (!mark[i]) continue;
            mark[i] = 0;
            dest_cache[i] = NULL_PART;
        }
        int32_t cap = deg < k ? deg : k;
        for (int32_t s = 0; s < cap; s++) tab[s][t] = make_int2(NULL_PART, 0);
        int32_t used = 0;
        for (eoff_t j = rs; j < re; j++) {
            part_t p = part[g.col[j]];
            int32_t po = p % cap;
            if (cap < k) {
                while (tab[po][t].x != NULL_PART && tab[po][t].x != p) po = (po + 1) % cap;
            }
            if (tab[po][t].x == NULL_PART) {
                tab[po][t].x = p;
                used++;
            }
            tab[po][t].y += g.ew[j];
        }
        int2* out = conn + conn_off[i];
        int32_t size = shrink_size(used, cap);
        if (size < cap) {
            for (int32_t s = 0; s < cap; s++) out[s] = make_int2(NULL_PART, 0);
            for (int32_t s = 0; s < cap; s++) {
                int2 e = tab[s][t];
                if (e.x == NULL_PART) continue;
                int32_t po = e.x % size;
                while (out[po].x != NULL_PART) po = (po + 1) % size;
                out[po] = e;
            }
        } else {
            for (int32_t s = 0; s < cap; s++) out[s] = tab[s][t];
        }
        conn_size[i] = size;
    }
}

// Inserts part p with weight w into an LDS table of `cap` slots (cap == k:
// direct; else probing); counts newly used slots.
__device__ inline void lds_table_add(int2* T, int32_t cap, int32_t k, part_t p, int32_t w, int32_t* used) {
    int32_t po = p % cap;
    if (cap == k) {
        if (T[po].x == NULL_PART && atomicCAS(&T[po].x, NULL_PART, p) == NULL_PART) atomicAdd(used, 1);
    } else {
        for (int32_t guard = 0; guard <= cap; guard++) {
            int32_t px = T[po].x;
            if (px == p) break;
            if (px == NULL_PART) {
                int32_t old = atomicCAS(&T[po].x, NULL_PART, p);
                if (old == NULL_PART) {
                    atomicAdd(used, 1);
                    break;
                }
                if (old == p) break;
            }
            po = (po + 1) % cap;
        }
    }
    atomicAdd(&T[po].y, w);
}

// Re-inserts the entries of T (cap slots) into T2 (size slots, linear probing).
__device__ inline void lds_table_shrink(const int2* T, int32_t cap, int2* T2, int32_t size, int start, int stride) {
    for (int32_t s = start; s < cap; s += stride) {
        int2 e = T[s];
        if (e.x == NULL_PART) continue;
        int32_t po = e.x % size;
        for (int32_t guard = 0; guard <= size; guard++) {
            if (T2[po].x == NULL_PART && atomicCAS(&T2[po].x, NULL_PART, e.x) == NULL_PART) {
                T2[po].y = e.y;
                break;
            }
            po = (po + 1) % size;
        }
    }
}

constexpr int WB_WAVES = 4;

__global__ void __launch_bounds__(WB_WAVES* WAVE)
    k_conn_build_wave(DGraph g, const part_t* part, int32_t k, const eoff_t* conn_off, int2* conn,
                      int32_t* conn_size, int32_t* mark, part_t* dest_cache, const vid_t* list, int32_t nlist,
                      const int32_t* gate) {
    if (gate && gate[0] == 0) return;
    // per wave two tables of tcap = min(k, WAVE_TABLE_MAX) slots, sized at launch
    extern __shared__ int2 tabs[];
    __shared__ int32_t used[WB_WAVES];
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int32_t tcap = k < WAVE_TABLE_MAX ? k : WAVE_TABLE_MAX;
    int2* Tw = tabs + (size_t)(2 * w) * tcap;
    int2* T2w = Tw + tcap;
    for (int32_t r = blockIdx.x * WB_WAVES + w; r < nlist; r += gridDim.x * WB_WAVES) {
        vid_t i = list[r];
        if (mark) {
            if (!mark[i]) continue;
        }
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        int32_t deg = (int32_t)(re - rs);
        int32_t cap = deg < k ? deg : k;
        for (int32_t s = lane; s < cap; s += WAVE) {
            Tw[s] = make_int2(NULL_PART, 0);
            T2w[s] = make_int2(NULL_PART, 0);
        }
        if (lane == 0) used[w] = 0;
        wave_sync();
        for (eoff_t j = rs + lane; j < re; j += WAVE) lds_table_add(Tw, cap, k, part[g.col[j]], g.ew[j], &used[w]);
        wave_sync();
        int32_t size = shrink_size(used[w], cap);
        int2* out = conn + conn_off[i];
        if (size < cap) {
            lds_table_shrink(Tw, cap, T2w, size, lane, WAVE);
            wave_sync();
            for (int32_t s = lane; s < cap; s += WAVE) out[s] = T2w[s];
        } else {
            for (int32_t s = lane; s < cap; s += WAVE) out[s] = Tw[s];
        }
        if (lane == 0) {
            conn_size[i] = size;
            if (mark) {
                mark[i] = 0;
                dest_cache[i] = NULL_PART;
            }
        }
        wave_sync();
    }
}

constexpr int BB_BLOCK = 512;

__global__ void __launch_bounds__(BB_BLOCK)
    k_conn_build_block(DGraph g, const part_t* part, int32_t k, const eoff_t* conn_off, int2* conn,
                       int32_t* conn_size, int32_t* mark, part_t* dest_cache, const vid_t* list, int32_t nlist,
                       const int32_t* gate) {
    if (gate && gate[0] == 0) return;
    // two tables of min(k, max degree) slots, sized at launch
    extern __shared__ int2 tabs[];
    const int32_t tcap = k;
    int2* T = tabs;
    int2* T2 = tabs + tcap;
    __shared__ int32_t used;
    for (int32_t r = blockIdx.x; r < nlist; r += gridDim.x) {
        vid_t i = list[r];
        if (mark && !mark[i]) continue;
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        int64_t deg = re - rs;
        int32_t cap = (int32_t)(deg < k ? deg : k);
        for (int32_t s = threadIdx.x; s < cap; s += BB_BLOCK) {
            T[s] = make_int2(NULL_PART, 0);
            T2[s] = make_int2(NULL_PART, 0);
        }
        if (threadIdx.x == 0) used = 0;
        __syncthreads();
        for (eoff_t j = rs + threadIdx.x; j < re; j += BB_BLOCK)
            lds_table_add(T, cap, k, part[g.col[j]], g.ew[j], &used);
        __syncthreads();
        int32_t size = shrink_size(used, cap);
        int2* out = conn + conn_off[i];
        if (size < cap) {
            lds_table_shrink(T, cap, T2, size, threadIdx.x, BB_BLOCK);
            __syncthreads();
            for (int32_t s = threadIdx.x; s < cap; s += BB_BLOCK) out[s] = T2[s];
        } else {
            for (int32_t s = threadIdx.x; s < cap; s += BB_BLOCK) out[s] = T[s];
        }
        if (threadIdx.x == 0) {
            conn_size[i] = size;
            if (mark) {
                mark[i] = 0;
                dest_cache[i] = NULL_PART;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- jet_lp
// jet_refiner.hpp:211-268: destination selection, viable-move filter.
// Each wave owns a contiguous chunk of CHW vertices (CHW / WAVE passes);
// candidates are staged in LDS and published with one atomic per workgroup
// of W waves.
constexpr int LS_WAVES = 8;
// Levels up to Tuning::lp_one_pass_n vertices run one pass per wave in
// workgroups of LS_WAVES waves (eight: 0.3 ms less per 128^3 stencil
// partition than four; sixteen is slower there). Larger levels run
// LS_MULTI_CHW vertices per wave in workgroups of LS_MULTI_WAVES waves: few
// enough workgroups appending to the candidate counters, and two passes per
// wave (measured on the 128^3 stencil: 12.1 us per launch on average,
// against 14.2 with eight passes in workgroups of four waves).
constexpr int LS_MULTI_CHW = 128;
constexpr int LS_MULTI_WAVES = 16;

template <int CHW, int W>
__global__ void __launch_bounds__(W* WAVE)
    k_lp_select(DGraph g, const part_t* part, const eoff_t* conn_off, const int2* conn, const int32_t* conn_size,
                part_t* dest_cache, int32_t* gain_p, int32_t* lock, double fr, part_t* dest, int32_t* pregain,
                vid_t* cand, int32_t* counters, vid_t nmax, int32_t* next_counters, unsigned long long* pc,
                BestCopy bc) {
    constexpr int NPASS = CHW / WAVE;
    __shared__ vid_t sbuf[W * CHW];
    __shared__ int32_t sh[2 * W + 2];
    __shared__ int64_t pred[W];
    // wave-private: the slots (u * WAVE + lane) of the wave's cache misses,
    // and each miss's new destination and gain
    __shared__ int16_t mslot[W][CHW];
    __shared__ part_t mbest[W][CHW];
    __shared__ int32_t mgain[W][CHW];
    int64_t n_miss = 0, n_slot = 0;
    StagedLists<W, CHW> st(sbuf);
    if (blockIdx.x == 0 && threadIdx.x < 8) next_counters[threadIdx.x] = 0;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = threadIdx.x / WAVE;
    const int64_t wbase = ((int64_t)blockIdx.x * W + wv) * CHW;
    const uint64_t lt = (1ull << lane) - 1ull;
    // 1. the cached destinations and locks of all passes, loaded together
    part_t cached[NPASS];
    int32_t locks[NPASS];
#pragma unroll
    for (int u = 0; u < NPASS; u++) {
        const int64_t i = wbase + u * WAVE + lane;
        cached[u] = i < g.n ? dest_cache[i] : NO_MOVE;
        locks[u] = i < g.n ? lock[i] : 0;
    }
    if (bc.part_out) {  // the previous step's best state, with the loads above
#pragma unroll
        for (int u = 0; u < NPASS; u++) {
            const int64_t i = wbase + u * WAVE + lane;
            if (i < g.n) bc.part_out[i] = part[i];
        }
        bc.sizes_block0();
    }
    // 2. the wave's misses, compacted, so that every lane evaluates one
    // table at a time instead of each pass waiting on its own few misses
    int32_t nm = 0;
#pragma unroll
    for (int u = 0; u < NPASS; u++) {
        const bool miss = cached[u] == NULL_PART;
        const uint64_t mk = __ballot(miss);
        if (miss) mslot[wv][nm + (int32_t)__popcll(mk & lt)] = (int16_t)(u * WAVE + lane);
        nm += (int32_t)__popcll(mk);
    }
    wave_sync();
    for (int32_t m = lane; m < nm; m += WAVE) {
        const int32_t slot = mslot[wv][m];
        const int64_t i = wbase + slot;
        const part_t p = part[i];
        part_t best = NO_MOVE;
        int32_t b_conn = 0, p_conn = 0;
        const int32_t sz = conn_size[i];
        n_miss++;
        n_slot += sz;
        for_slots(conn + conn_off[i], sz, [&](int2 e) {
            if (e.y > b_conn && e.x != p) {
                best = e.x;
                b_conn = e.y;
            } else if (e.y == b_conn && b_conn > 0 && e.x != p && e.x < best) {
                best = e.x;
            } else if (e.y > 0 && e.x == p) {
                p_conn = e.y;
            }
        });
        int32_t gsv = 0;
        if (best != NO_MOVE) {
            if (b_conn >= p_conn || ((double)(p_conn - b_conn) < floor(fr * (double)p_conn))) gsv = b_conn - p_conn;
            else best = NO_MOVE;
        }
        gain_p[i] = gsv;
        dest_cache[i] = best;
        mbest[wv][slot] = best;
        mgain[wv][slot] = gsv;
    }
    wave_sync();
    // 3. destinations; a vertex that is not a candidate this iteration gets
    // NO_MOVE: the afterburner skips such neighbours before reading their
    // pre-gain, which is then only written for candidates. The candidates'
    // gains and row offsets of all passes are loaded together.
    bool is_cand[NPASS];
    int32_t gain[NPASS];
    eoff_t r0[NPASS], r1[NPASS];
#pragma unroll
    for (int u = 0; u < NPASS; u++) {
        const int64_t i = wbase + u * WAVE + lane;
        const bool miss = cached[u] == NULL_PART;
        const part_t best = miss ? mbest[wv][u * WAVE + lane] : cached[u];
        is_cand[u] = i < g.n && best != NO_MOVE && locks[u] == 0;
        if (i < g.n) {
            dest[i] = is_cand[u] ? best : NO_MOVE;
            if (!is_cand[u] && locks[u]) lock[i] = 0;
        }
        gain[u] = 0;
        r0[u] = r1[u] = 0;
        if (is_cand[u]) {
            gain[u] = miss ? mgain[wv][u * WAVE + lane] : gain_p[i];
            r0[u] = g.rowptr[i];
            r1[u] = g.rowptr[i + 1];
        }
    }
#pragma unroll
    for (int u = 0; u < NPASS; u++) {
        const int64_t i = wbase + u * WAVE + lane;
        const eoff_t deg = r1[u] - r0[u];
        if (is_cand[u]) pregain[i] = gain[u];
        st.push(is_cand[u] && deg <= BIG_DEG, is_cand[u] && deg > BIG_DEG, (vid_t)i);
    }
    st.flush(cand, &counters[0], &counters[1], nmax, sh);
    if (pc) {  // counting runs only: evaluated vertices and table slots read
        int64_t v2[2] = {n_miss, n_slot};
        for (int q = 0; q < 2; q++) {
            int64_t v = block_reduce_sum<int64_t, W * WAVE>(v2[q], pred);
            if (threadIdx.x == 0 && v) atomicAdd(&pc[q], (unsigned long long)v);
        }
    }
}

// 16-wave workgroups: each workgroup publishes its moves with one atomic on
// the move counters, and fewer workgroups contend for them (measured 0.9 ms
// less per 128^3 stencil partition than four-wave workgroups)
constexpr int AB_WAVES = 16;
constexpr int AB_CHW = 256;

template <int G>
__global__ void __launch_bounds__(AB_WAVES* WAVE)
    k_afterburner_small(DGraph g, const vid_t* cand, const int32_t* counters, const part_t* part, const part_t* dest,
                        const int32_t* pregain, int32_t* lock, vid_t* moves, int32_t* mcounters, vid_t nmax,
                        int32_t* stamp, int32_t gen) {
    __shared__ vid_t sbuf[AB_WAVES * AB_CHW];
    __shared__ int32_t sh[2 * AB_WAVES + 2];
    StagedLists<AB_WAVES, AB_CHW> st(sbuf);
    const int lane = threadIdx.x & (WAVE - 1);
    const int sub = lane & (G - 1);
    // wave w takes candidates w * WAVE/G + [0, WAVE/G), then every tw-th such
    // group (tw = waves in the grid): the mapping does not depend on the
    // count, so a lane's first candidate is loaded together with it. The
    // host launches tw * AB_CHW >= n lanes' worth of waves, so no wave
    // stages more than AB_CHW moves (AB_CHW * G / WAVE rounds).
    const int64_t tw = (int64_t)gridDim.x * AB_WAVES;
    const int64_t wfirst = ((int64_t)blockIdx.x * AB_WAVES + threadIdx.x / WAVE) * (WAVE / G);
    const int64_t x0 = wfirst + lane / G;
    const vid_t c0 = x0 < nmax ? cand[x0] : 0;
    const int32_t nc = counters[0];
    int32_t round = 0;
    for (int64_t xw = wfirst; xw < nc && round < AB_CHW * G / WAVE; xw += tw * (WAVE / G), round++) {
        const int64_t x = xw + lane / G;
        bool keep = false;
        vid_t i = 0;
        if (x < nc) {
            i = round == 0 ? c0 : cand[x];
            part_t p = part[i], best = dest[i];
            int32_t ig = pregain[i];
            int64_t change = 0;
            for (eoff_t j = g.rowptr[i] + sub; j < g.rowptr[i + 1]; j += G)
                change += afterburn_edge(g.col[j], i, ig, g.ew[j], p, best, part, dest, pregain);
#pragma unroll
            for (int d = G / 2; d >= 1; d >>= 1) change += __shfl_xor(change, d, G);
            keep = (int64_t)ig + change >= 0;
            if (keep && sub == 0) lock[i] = 1;
        }
        st.push(keep && sub == 0, false, i);
    }
    st.flush_stamped(moves, &mcounters[2], &mcounters[3], nmax, sh, stamp, gen);
}

__global__ void k_afterburner_big(DGraph g, const vid_t* cand, const int32_t* counters, const part_t* part,
                                  const part_t* dest, const int32_t* pregain, int32_t* lock, vid_t* moves,
                                  int32_t* mcounters, vid_t nmax, int32_t* stamp, int32_t gen) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t wstride = ((int64_t)gridDim.x * blockDim.x) / WAVE;
    const int64_t x0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
    const vid_t c0 = x0 < nmax ? cand[nmax - 1 - x0] : 0;  // loaded together with the count
    const int32_t nb = counters[1];
    for (int64_t x = x0; x < nb; x += wstride) {
        const vid_t i = x == x0 ? c0 : cand[nmax - 1 - x];
        part_t p = part[i], best = dest[i];
        int32_t ig = pregain[i];
        int64_t change = 0;
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        for (eoff_t j = rs + lane; j < re; j += WAVE)
            change += afterburn_edge(g.col[j], i, ig, g.ew[j], p, best, part, dest, pregain);
        change = wave_reduce_sum(change);
        if (lane == 0 && (int64_t)ig + change >= 0) {
            lock[i] = 1;
            stamp[i] = gen;
            if (re - rs > BIG_DEG) moves[nmax - 1 - atomicAdd(&mcounters[3], 1)] = i;
            else moves[atomicAdd(&mcounters[2], 1)] = i;
        }
    }
}

// ---------------------------------------------------------------- perform_moves
// perform_moves' choice between update_small (push each moved vertex's
// contribution into its neighbours' tables) and update_large (rebuild the
// tables of every neighbour of a moved vertex): rebuild when more than n/10
// vertices move (jet_refiner.hpp:933). The two do not leave the same tables:
// pushing keeps a part whose connectivity falls to 0 in a table of k slots
// (jet_refiner.hpp:806-810) and the rebalancers read such entries, so the
// choice follows the reference exactly.
__device__ inline bool update_by_rebuild(const int32_t* mc, vid_t n) { return mc[2] + mc[3] > n / 10; }

// The cut change and the largest part (jet_refiner.hpp:937-948). For a moved
// vertex v the reference adds conn(v, new part) - conn(v, old part) from its
// table before the updates and again after them; each is the sum over v's
// edges (v, u) of w * ([part(u) == new part of v] - [part(u) == old part of v]),
// with the parts before and after the moves, so the first pass computes both
// from the parts and needs no table. It adds each workgroup's sum into one of
// PUB_GROUPS partials (workgroup b into partial b mod PUB_GROUPS), 128 bytes
// apart, so that no line takes more than a sixteenth of the grid's atomics.
// Workgroup 0 of the second pass folds them, takes the largest part and
// writes (cut change, largest part, deferred, rebuild) to the host mailbox. A
// deferred step reports only the deferral.
constexpr int PUB_GROUPS = 16;
constexpr int PUB_STRIDE = 16;  // int64 words: 128 bytes

__device__ inline void add_cut_part2(int64_t* partials, int64_t c2, int64_t* red) {
    c2 = block_reduce_sum<int64_t, 256>(c2, red);
    if (threadIdx.x == 0 && c2)
        atomicAdd((unsigned long long*)&partials[(int64_t)(blockIdx.x % PUB_GROUPS) * PUB_STRIDE],
                  (unsigned long long)c2);
}

// Run by one workgroup (256 threads) after the first pass: folds the
// partials, takes the largest part and writes (cut change, largest part,
// deferred, rebuild) to the host mailbox, then the sequence word the host
// waits for, and clears the accumulators. One workgroup needs no completion
// ticket (a ticket in a many-workgroup kernel needs an agent-scope release
// fence per workgroup, which writes back L2), and the kernel boundary before
// it orders the other kernels' atomics.
// The state a rebalancing step starts from, besides zeroed bucket sums: the
// undersized parts in increasing id (weak, jet_refiner.hpp:578-591) or the
// largest receiver capacity and cleared eviction adjustments (strong,
// :440-448), from the part sizes before the step. One workgroup of 256.
__device__ inline void rebalance_prep_block(const int64_t* sizes, int32_t k, int64_t max_dest, bool strong,
                                            part_t* undersized, int32_t* misc, int64_t* acc, int64_t* evict_adjust,
                                            int64_t* red, int32_t* ws) {
    if (strong) {
        int64_t mx = INT64_MIN;
        for (int t = threadIdx.x; t < k; t += blockDim.x) {
            int64_t sz = sizes[t];
            if (sz < max_dest && max_dest - sz > mx) mx = max_dest - sz;
            evict_adjust[t] = 0;
        }
        mx = block_reduce_max<int64_t, 256>(mx, red);
        if (threadIdx.x == 0) acc[3] = mx;
    } else {
        int32_t base = 0;
        for (int32_t p0 = 0; p0 < k; p0 += 256) {
            int32_t p = p0 + threadIdx.x;
            int32_t f = (p < k && sizes[p] < max_dest) ? 1 : 0;
            int32_t tot;
            int32_t ex = block_excl_scan<int32_t, 256>(f, tot, ws);
            if (f) undersized[base + ex] = p;
            base += tot;
        }
        if (threadIdx.x == 0) misc[0] = base;
    }
}

// What the second update pass prepares for the next step in case it is a
// rebalancing step of the given kind (buckets == nullptr: nothing): its
// bucket sums zeroed (nb entries) and rebalance_prep_block's state, so that
// the step needs no reset launch.
struct RebalPrep {
    int64_t* buckets;
    int64_t nb;
    int64_t max_dest;
    int32_t strong;
    part_t* undersized;
    int32_t* misc;
    int64_t* acc;
    int64_t* evict_adjust;
};

// The first update pass's workgroups add their part-size deltas into one of
// `ncopies` copies of a k-entry array (workgroup b into copy b mod ncopies),
// so that fewer workgroups' atomics meet on one address; the second pass's
// publishing workgroup folds the copies into the sizes and clears them.
constexpr int SIZE_COPIES_MAX = 16;

struct Publish {
    int64_t* partials;
    int64_t* sizes;
    int32_t k;
    int64_t* host_out;
    const int32_t* mc;
    int64_t seq;
    unsigned long long* copies;  // ncopies * k deltas, zero between steps
    int32_t ncopies;
};

__device__ inline void fold_size_copies(const Publish& pb) {
    if (pb.ncopies <= 0) return;
    for (int t = threadIdx.x; t < pb.k; t += blockDim.x) {
        // every copy's load in flight at once, then the clears
        unsigned long long v[SIZE_COPIES_MAX];
#pragma unroll
        for (int c = 0; c < SIZE_COPIES_MAX; c++) v[c] = c < pb.ncopies ? pb.copies[(int64_t)c * pb.k + t] : 0;
        unsigned long long d = 0;
#pragma unroll
        for (int c = 0; c < SIZE_COPIES_MAX; c++) {
            d += v[c];
            if (v[c]) pb.copies[(int64_t)c * pb.k + t] = 0;
        }
        if (d) pb.sizes[t] += (int64_t)d;
    }
    __syncthreads();
}

__device__ inline void publish_step(const Publish& pb, bool deferred, int64_t* red) {
    if (deferred) {
        if (threadIdx.x == 0) {
            pb.host_out[2] = 1;
            __threadfence_system();
            pb.host_out[5] = pb.seq;  // the host waits for this word
            __threadfence_system();
        }
        return;
    }
    int64_t mx = INT64_MIN;
    for (int t = threadIdx.x; t < pb.k; t += blockDim.x) mx = pb.sizes[t] > mx ? pb.sizes[t] : mx;
    mx = block_reduce_max<int64_t, 256>(mx, red);
    int64_t c2 = threadIdx.x < PUB_GROUPS ? pb.partials[(int64_t)threadIdx.x * PUB_STRIDE] : 0;
    c2 = block_reduce_sum<int64_t, 256>(c2, red);
    if (threadIdx.x < PUB_GROUPS) pb.partials[(int64_t)threadIdx.x * PUB_STRIDE] = 0;
    if (threadIdx.x == 0) {
        pb.host_out[0] = c2;
        pb.host_out[1] = mx;
        pb.host_out[2] = 0;
        pb.host_out[3] = pb.mc[6];  // the tables are to be rebuilt (update_large)
        __threadfence_system();
        pb.host_out[5] = pb.seq;  // written after the results: the host waits for it
        __threadfence_system();
    }
}

// One move's record, written by the first pass for the second: the vertex,
// its new part and its row, so that the second pass's chain starts at the
// record (no part, destination or row-offset loads).
struct MoveRec {
    eoff_t rs;
    int32_t deg;
    vid_t v;
    part_t pnew;
    int32_t unused;
};

// One move in either pass, or one slice of it: the edges rs + sub, rs + sub +
// stride, ... before re.
template <bool ADD>
__device__ inline void update_move(const DGraph& g, vid_t i, part_t pold, part_t pnew, eoff_t rs, eoff_t re, int sub,
                                   int stride,
                                   bool rebuild, const part_t* part, const part_t* dest, const eoff_t* conn_off,
                                   int2* conn, int32_t* conn_size, part_t* dest_cache, int32_t k, int32_t* err,
                                   int32_t* mark, const uint8_t* hub, uint8_t* hub_dirty, const int32_t* stamp,
                                   int32_t gen, int64_t& cc) {
    for (eoff_t j = rs + sub; j < re; j += stride) {
        const vid_t v = g.col[j];
        const int32_t w = g.ew[j];
        if (!ADD) {
            // loaded together: the neighbour's parts and table
            const part_t pv = part[v], dv = dest[v];
            const bool sv = stamp[v] == gen;
            if (rebuild) {
                const part_t nv = sv ? dv : pv;
                cc += (int64_t)w *
                      ((pv == pnew ? 1 : 0) - (pv == pold ? 1 : 0) + (nv == pnew ? 1 : 0) - (nv == pold ? 1 : 0));
                mark[v] = 1;
                continue;
            }
            const uint8_t hv = hub[v];
            const eoff_t co = conn_off[v];
            const int32_t cs = ld_rx(&conn_size[v]);
            const part_t nv = sv ? dv : pv;
            cc += (int64_t)w *
                  ((pv == pnew ? 1 : 0) - (pv == pold ? 1 : 0) + (nv == pnew ? 1 : 0) - (nv == pold ? 1 : 0));
            if (hv) {  // updated by k_update_hubs
                hub_dirty[v] = 1;
                continue;
            }
            conn_sub(conn + co, cs, pold, w, k, err);
        } else {
            const uint8_t hv = hub[v];
            const eoff_t co = conn_off[v], ce = conn_off[v + 1];
            if (hv) continue;  // updated by k_update_hubs
            dest_cache[v] = NULL_PART;
            conn_add(conn + co, &conn_size[v], (int32_t)(ce - co), pnew, w, err);
        }
    }
}

// The moves of a step, in two passes (jet_refiner.hpp:905-955 and
// update_small/update_large, :785-884). A move v goes from part[v] to dest[v];
// the producers of the move list stamp each listed vertex with the step's
// generation, so the new part of any vertex u is dest[u] when stamp[u] == gen
// and part[u] otherwise, before part[] changes.
// ADD == false (first pass): part-size deltas of the moves, the moved
// vertices' cached destinations cleared, the whole cut change (for each edge
// (v, u) of a moved v: w * ([p(u) == new] - [p(u) == old]) with the parts
// before and after the step, the reference's table lookups before and after
// the updates), and the old part's connectivity subtracted from each
// neighbour's table, or, when the tables are rebuilt instead (update_large),
// the neighbours marked; each move's record for the second pass.
// ADD == true (second pass): part[v] = new part and the new part's
// connectivity added (unless rebuilding); workgroup 0 only publishes the
// step, so that its chain of loads does not add to a chain of move updates.
// Each kernel loads the deferral flag, the move counts and its first list
// entry together. G lanes stride the edges of one move.
template <bool ADD, int G>
__global__ void __launch_bounds__(256)
    k_update_small(DGraph g, const vid_t* moves, MoveRec* rec, int32_t* mc, vid_t nmax, part_t* part,
                   const part_t* dest, const eoff_t* conn_off, int2* conn, int32_t* conn_size, part_t* dest_cache,
                   int32_t k, int32_t* err, int32_t* mark, const uint8_t* hub, uint8_t* hub_dirty,
                   const int32_t* deferred, const int32_t* stamp, int32_t gen, unsigned long long* sizes,
                   int64_t* partials, Publish pb, RebalPrep prep) {
    __shared__ int64_t red[256 / WAVE];
    __shared__ int32_t ws[ADD ? 256 / WAVE + 1 : 1];
    __shared__ unsigned long long ls[ADD ? 1 : LDS_K];
    const int sub = threadIdx.x & (G - 1);
    // the second pass's move workgroups are 1 .. gridDim.x - 1
    const int64_t blk = ADD ? (int64_t)blockIdx.x - 1 : (int64_t)blockIdx.x;
    const int64_t nblk = ADD ? (int64_t)gridDim.x - 1 : (int64_t)gridDim.x;
    const int64_t gstride = (nblk * blockDim.x) / G;
    const int64_t x0 = ADD && blockIdx.x == 0 ? nmax : (blk * blockDim.x + threadIdx.x) / G;
    const bool have0 = x0 < nmax;
    const vid_t i0 = !ADD && have0 ? moves[x0] : 0;
    MoveRec r0{};
    if (ADD && have0) r0 = rec[x0];
    const int32_t dfr = *deferred;
    const int32_t ns = mc[2], nbig = mc[3];
    // the second pass's workgroup 0 publishes the step (the first pass, which
    // completes the results, has finished)
    if (ADD && blockIdx.x == 0) {
        // the first pass's size deltas (every copy complete: that pass has
        // finished); no other kernel of this pass touches the sizes
        fold_size_copies(pb);
        publish_step(pb, dfr != 0, red);
        // then the next rebalancing step's start (its kernels run after this one)
        if (!dfr && prep.buckets)
            rebalance_prep_block(pb.sizes, pb.k, prep.max_dest, prep.strong != 0, prep.undersized, prep.misc,
                                 prep.acc, prep.evict_adjust, red, ws);
        return;
    }
    if (dfr) return;
    if (ADD && prep.buckets) {
        // the move workgroups zero the next rebalancing step's bucket sums (the
        // current step's readers have all run before the first pass)
        for (int64_t i = (blk * blockDim.x) + threadIdx.x; i < prep.nb; i += nblk * blockDim.x) prep.buckets[i] = 0;
    }
    const bool rebuild = ns + nbig > g.n / 10;  // update_by_rebuild
    if (!ADD && rebuild && blockIdx.x == 0 && threadIdx.x == 0) mc[6] = 1;  // the tables are to be rebuilt
    const bool lds = !ADD && k <= LDS_K;
    if (lds) {
        for (int t = threadIdx.x; t < k; t += blockDim.x) ls[t] = 0;
        __syncthreads();
    }
    int64_t cc = 0;
    for (int64_t x = x0; x < ns; x += gstride) {
        if (ADD) {
            const MoveRec r = x == x0 ? r0 : rec[x];
            if (sub == 0) part[r.v] = r.pnew;
            if (rebuild) continue;
            update_move<true>(g, r.v, 0, r.pnew, r.rs, r.rs + r.deg, sub, G, false, part, dest, conn_off, conn,
                                 conn_size, dest_cache, k, err, mark, hub, hub_dirty, stamp, gen, cc);
        } else {
            const vid_t i = x == x0 ? i0 : moves[x];
            const part_t pold = part[i], pnew = dest[i];
            const eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
            if (sub == 0) {
                dest_cache[i] = NULL_PART;
                rec[x] = MoveRec{rs, (int32_t)(re - rs), i, pnew, 0};
                const unsigned long long w = (unsigned long long)g.vw[i];
                if (lds) {
                    atomicAdd(&ls[pold], (unsigned long long)(-(long long)w));
                    atomicAdd(&ls[pnew], w);
                } else {
                    atomicAdd(&sizes[pold], (unsigned long long)(-(long long)w));
                    atomicAdd(&sizes[pnew], w);
                }
            }
            update_move<false>(g, i, pold, pnew, rs, re, sub, G, rebuild, part, dest, conn_off, conn, conn_size,
                                  dest_cache, k, err, mark, hub, hub_dirty, stamp, gen, cc);
        }
    }
    if (!ADD) {
        add_cut_part2(partials, cc, red);
        if (lds) {
            __syncthreads();
            unsigned long long* dst =
                pb.ncopies > 0 ? pb.copies + (int64_t)(blockIdx.x % (uint32_t)pb.ncopies) * k : sizes;
            for (int t = threadIdx.x; t < k; t += blockDim.x)
                if (ls[t]) atomicAdd(&dst[t], ls[t]);
        }
    }
}

// The same for the moves of degree > BIG_DEG (the back of the list), one wave
// per move.
template <bool ADD>
__global__ void __launch_bounds__(256)
    k_update_big(DGraph g, const vid_t* moves, MoveRec* rec, const int32_t* mc, vid_t nmax, part_t* part,
                 const part_t* dest, const eoff_t* conn_off, int2* conn, int32_t* conn_size, part_t* dest_cache,
                 int32_t k, int32_t* err, int32_t* mark, const uint8_t* hub, uint8_t* hub_dirty,
                 const int32_t* deferred, const int32_t* stamp, int32_t gen, unsigned long long* sizes,
                 int64_t* partials, int32_t slices) {
    __shared__ int64_t red[256 / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    // units: (move x, slice q) for q < slices; slice q of a move takes its
    // edges q * WAVE + lane, stride slices * WAVE, so a hub's edges are
    // spread over `slices` waves; slice 0 does the move's bookkeeping
    const int64_t wstride = ((int64_t)gridDim.x * blockDim.x) / WAVE;
    const int64_t u0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
    const int64_t x0 = u0 / slices;
    const bool have0 = x0 < nmax;
    const vid_t i0 = !ADD && have0 ? moves[nmax - 1 - x0] : 0;
    MoveRec r0{};
    if (ADD && have0) r0 = rec[nmax - 1 - x0];
    const int32_t dfr = *deferred;
    const int32_t ns = mc[2], nb = mc[3];
    if (dfr) return;
    const bool rebuild = ns + nb > g.n / 10;  // update_by_rebuild
    const int STRIDE = slices * WAVE;
    int64_t cc = 0;
    for (int64_t u = u0; u < (int64_t)nb * slices; u += wstride) {
        const int64_t x = u / slices;
        const int q = (int)(u % slices);
        const int64_t at = nmax - 1 - x;
        if (ADD) {
            const MoveRec r = x == x0 ? r0 : rec[at];
            if (q == 0 && lane == 0) part[r.v] = r.pnew;
            if (rebuild) continue;
            update_move<true>(g, r.v, 0, r.pnew, r.rs, r.rs + r.deg, q * WAVE + lane, STRIDE, false, part, dest,
                              conn_off, conn, conn_size, dest_cache, k, err, mark, hub, hub_dirty, stamp, gen, cc);
        } else {
            const vid_t i = x == x0 ? i0 : moves[at];
            const part_t pold = part[i], pnew = dest[i];
            const eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
            if (q == 0 && lane == 0) {
                dest_cache[i] = NULL_PART;
                rec[at] = MoveRec{rs, (int32_t)(re - rs), i, pnew, 0};
                const unsigned long long w = (unsigned long long)g.vw[i];
                atomicAdd(&sizes[pold], (unsigned long long)(-(long long)w));
                atomicAdd(&sizes[pnew], w);
            }
            update_move<false>(g, i, pold, pnew, rs, re, q * WAVE + lane, STRIDE, rebuild, part, dest, conn_off,
                               conn, conn_size, dest_cache, k, err, mark, hub, hub_dirty, stamp, gen, cc);
        }
    }
    if (!ADD) add_cut_part2(partials, cc, red);
}

// update_small for rows of degree > BIG_DEG ("hubs": the workgroup-built tier
// and the pull_mid rows), one workgroup per row with a moved neighbour: the
// row's moved neighbours are found by a scan of its adjacency and their
// weights summed per part in LDS, then each part's total is applied to the
// table once, with the subtraction phase before the
// addition phase as for the other rows. The resulting table equals the one
// per-edge pushes leave: a part is reclaimed when its connectivity reaches 0
// in a table of fewer than k slots (jet_refiner.hpp:806-810), new parts take
// free or reclaimed slots and then overflow slots (:828-884). Pushing instead
// would land every moved neighbour's atomics on the same few hub slots.
template <bool ADD, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    k_update_hubs(DGraph g, const vid_t* list, int32_t nlist, const int32_t* mc, const int32_t* stamp, int32_t gen,
                  uint8_t* hub_dirty, const part_t* part, const part_t* dest, const eoff_t* conn_off, int2* conn,
                  int32_t* conn_size, part_t* dest_cache, int32_t k, int32_t* err, const int32_t* deferred) {
    if (*deferred || update_by_rebuild(mc, g.n)) return;
    extern __shared__ int32_t delta[];  // k
    for (int32_t r = blockIdx.x; r < nlist; r += gridDim.x) {
        const vid_t u = list[r];
        if (!hub_dirty[u]) continue;
        for (int32_t p = threadIdx.x; p < k; p += BLOCK) delta[p] = 0;
        __syncthreads();
        for (eoff_t j = g.rowptr[u] + threadIdx.x; j < g.rowptr[u + 1]; j += BLOCK) {
            vid_t v = g.col[j];
            // between the passes part[v] is still the old part; the new one is dest[v]
            if (stamp[v] == gen) atomicAdd(&delta[ADD ? dest[v] : part[v]], g.ew[j]);
        }
        __syncthreads();
        int2* tab = conn + conn_off[u];
        const int32_t cap = (int32_t)(conn_off[u + 1] - conn_off[u]);
        const int32_t vsz = conn_size[u];
        for (int32_t p = threadIdx.x; p < k; p += BLOCK) {
            const int32_t w = delta[p];
            if (w == 0) continue;
            if (ADD) conn_add(tab, &conn_size[u], cap, p, w, err);
            else conn_sub(tab, vsz, p, w, k, err);
        }
        if (ADD && threadIdx.x == 0) {
            dest_cache[u] = NULL_PART;
            hub_dirty[u] = 0;
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------- rebalancing
// Rebalancer minibuckets (jet_refiner.hpp:359-412, :449-481, :632-643).
// Bucket sums are order-free, so they are accumulated with atomics and
// scanned. A vertex is evicted when pos + off[b] - off[begin_p] < limit_p,
// where pos is the weight of same-bucket candidates with smaller id. Whole
// buckets below a part's limit are decided without pos; only the members of
// buckets that straddle a limit (or, for the strong rebalancer, every bucket
// below it, whose scores feed the destination chunks) are ranked by id.

// Zeroes the bucket arrays and the next iteration's counters; workgroup 0
// also builds the undersized-part list (weak, jet_refiner.hpp:578-591) or
// the largest receiver capacity (strong, :440-448) and clears the eviction
// adjustments.
__global__ void __launch_bounds__(256)
    k_reset_buckets(int64_t* buckets, int64_t nb, int32_t* next_counters, const int64_t* sizes,
                    int32_t k, int64_t max_dest, int strong, part_t* undersized, int32_t* misc, int64_t* acc,
                    int64_t* evict_adjust) {
    FOR_THREAD(i, nb) buckets[i] = 0;
    if (blockIdx.x != 0) return;
    if (threadIdx.x < 8) next_counters[threadIdx.x] = 0;
    __shared__ int32_t ws[256 / WAVE + 1];
    __shared__ int64_t red[256 / WAVE];
    rebalance_prep_block(sizes, k, max_dest, strong != 0, undersized, misc, acc, evict_adjust, red, ws);
}

// Per-workgroup LDS map of bucket sums (a workgroup takes a contiguous range
// of vertices, and consecutive vertices share few buckets); overflow goes
// straight to global atomics.
constexpr int BH_SLOTS = 1024;

struct BucketMap {
    int32_t* key;
    unsigned long long* val;
    __device__ void init() {
        for (int t = threadIdx.x; t < BH_SLOTS; t += blockDim.x) {
            key[t] = -1;
            val[t] = 0;
        }
        __syncthreads();
    }
    __device__ void add(int32_t b, unsigned long long w, unsigned long long* global) {
        uint32_t h = ((uint32_t)b * 0x9E3779B1u) >> 22;
        int32_t slot = -1;
        for (int probe = 0; probe < 16; probe++) {
            int32_t k = key[h];
            if (k == -1) {
                int32_t old = atomicCAS(&key[h], -1, b);
                if (old == -1 || old == b) {
                    slot = (int32_t)h;
                    break;
                }
            } else if (k == b) {
                slot = (int32_t)h;
                break;
            }
            h = (h + 1) & (BH_SLOTS - 1);
        }
        // distinct scopes keep the LDS add and the global fallback apart (one
        // flat atomic for both would serve the common LDS case slowly)
        if (slot >= 0) __hip_atomic_fetch_add(&val[slot], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else atomicAdd(&global[b], w);
    }
    __device__ void flush(unsigned long long* global) {
        __syncthreads();
        for (int t = threadIdx.x; t < BH_SLOTS; t += blockDim.x)
            if (key[t] != -1) atomicAdd(&global[key[t]], val[t]);
    }
};

// The select kernels' workgroups flush their bucket maps into `copies`
// copies of the bucket array (workgroup b into copy b mod copies), so that
// fewer workgroups' atomics meet on one line of it; k_sum_bucket_copies folds
// copies 1.. into copy 0 (the bucket array itself) and clears them.
constexpr int BUCKET_COPIES_MAX = 16;
constexpr int64_t BUCKET_COPIES_BYTES = 64ll << 20;

__device__ inline unsigned long long* bucket_copy(unsigned long long* buckets, unsigned long long* extra,
                                                  int32_t copies, int64_t stride) {
    const int32_t c = (int32_t)(blockIdx.x % (uint32_t)copies);
    return c == 0 ? buckets : extra + (int64_t)(c - 1) * stride;
}

__global__ void k_sum_bucket_copies(int64_t* buckets, int64_t* extra, int32_t copies, int64_t nb, int64_t stride) {
    FOR_THREAD(j, nb) {
        int64_t sum = buckets[j];
        for (int32_t c0 = 0; c0 < copies - 1; c0 += 8) {
            int64_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = c0 + u < copies - 1 ? extra[(int64_t)(c0 + u) * stride + j] : 0;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                sum += v[u];
                if (v[u]) extra[(int64_t)(c0 + u) * stride + j] = 0;
            }
        }
        buckets[j] = sum;
    }
}

// Part sizes of the select kernels: staged in LDS when k <= SEL_LDS_K (each
// table slot's destination check is then an LDS read, not a dependent global
// load), read from global memory otherwise.
constexpr int SEL_LDS_K = 1024;

constexpr int SEL_BLOCK_MAX = 512;  // select kernels: 128 threads, 512 on large levels

template <bool LDS>
struct PartSizes {
    const int64_t* global;
    int64_t* lds;
    __device__ void stage(int32_t k) {
        if (LDS)
            for (int32_t t = threadIdx.x; t < k; t += blockDim.x) lds[t] = global[t];
    }
    __device__ int64_t operator[](int32_t p) const { return LDS ? lds[p] : global[p]; }
};

// jet_refiner.hpp:592-643: destination and bucket of each vertex of an
// oversized part.
template <bool LDS>
__global__ void __launch_bounds__(SEL_BLOCK_MAX)
    k_weak_select(DGraph g, const part_t* part, const int64_t* sizes, int32_t k, int64_t size_max, int32_t opt,
                  int64_t max_dest, const part_t* undersized, const int32_t* misc, const eoff_t* conn_off,
                  const int2* conn, const int32_t* conn_size, int32_t sections, part_t* dest, int32_t* bid,
                  unsigned long long* buckets_base, unsigned long long* extra, int32_t copies, int64_t stride,
                  unsigned long long* pc, int32_t* next_counters, BestCopy bc) {
    if (blockIdx.x == 0 && threadIdx.x < 8) next_counters[threadIdx.x] = 0;
    bc.sizes_block0();
    unsigned long long* buckets = bucket_copy(buckets_base, extra, copies, stride);
    __shared__ int32_t bk[BH_SLOTS];
    __shared__ unsigned long long bv[BH_SLOTS];
    __shared__ int64_t ssz[LDS ? SEL_LDS_K : 1];
    PartSizes<LDS> ps{sizes, ssz};
    ps.stage(k);
    BucketMap bm{bk, bv};
    bm.init();
    const int32_t tu = misc[0];
    int64_t n_eval = 0, n_slot = 0;
    // one contiguous range per workgroup (as FOR_BLOCK_RANGE), WS_V of a
    // thread's vertices per pass with their part and table-header loads
    // issued together: the dependent chain part -> header -> slots is paid
    // once per pass, not once per vertex
    constexpr int WS_V = 4;
    const int64_t per_block = ((g.n + gridDim.x - 1) / gridDim.x + blockDim.x - 1) / blockDim.x * blockDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per_block;
    const int64_t hi = min((int64_t)g.n, lo + per_block);
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)WS_V * blockDim.x) {
        part_t p[WS_V], spare[WS_V];
        wgt_t w[WS_V];
        eoff_t off[WS_V];
        int32_t sz[WS_V];
        int64_t sp[WS_V];
        bool ev[WS_V];
#pragma unroll
        for (int u = 0; u < WS_V; u++) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            p[u] = i < hi ? part[i] : 0;
            if (bc.part_out && i < hi) bc.part_out[i] = p[u];
        }
#pragma unroll
        for (int u = 0; u < WS_V; u++) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            sp[u] = ps[p[u]];
            const bool over = i < hi && sp[u] > size_max;
            w[u] = over ? g.vw[i] : 0;
            off[u] = over ? conn_off[i] : 0;
            sz[u] = over ? conn_size[i] : 0;
            spare[u] = over && tu > 0 ? undersized[(uint32_t)i % (uint32_t)tu] : p[u];
            ev[u] = over && (double)w[u] < 1.5 * (double)(sp[u] - opt);
        }
#pragma unroll
        for (int u = 0; u < WS_V; u++) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            if (i >= hi) continue;
            part_t d = p[u];
            int32_t b = -1;
            if (ev[u]) {
                n_eval++;
                n_slot += sz[u];
                part_t best = p[u];
                int32_t gain = 0, p_gain = 0;
                for_slots(conn + off[u], sz[u], [&](int2 e) {
                    if (e.x > NULL_PART && ps[e.x] < max_dest) {
                        if (e.y > gain) {
                            best = e.x;
                            gain = e.y;
                        } else if (e.y == gain && gain > 0 && e.x < best) {
                            best = e.x;
                        }
                    }
                    if (e.x == p[u]) p_gain = e.y;
                });
                int64_t sv = 0;
                if (gain > 0) {
                    d = best;
                    sv = (int64_t)gain - p_gain;
                } else if (tu > 0) {
                    d = spare[u];
                    sv = -(int64_t)p_gain;
                }
                if (d != p[u]) {
                    int32_t gt = gain_bucket_fast(sv, w[u]);
                    b = (MAX_BUCKETS * p[u] + gt) * sections + (int32_t)((uint32_t)i % (uint32_t)sections);
                    bm.add(b, (unsigned long long)w[u], buckets);
                }
            }
            dest[i] = d;
            bid[i] = b;
        }
    }
    bm.flush(buckets);
    if (pc) {  // counting runs only: evaluated vertices and table slots read
        __shared__ int64_t red[SEL_BLOCK_MAX / WAVE];
        int64_t v2[2] = {n_eval, n_slot};
        for (int q = 0; q < 2; q++) {
            const int64_t v = blockDim.x == SEL_BLOCK_MAX ? block_reduce_sum<int64_t, SEL_BLOCK_MAX>(v2[q], red)
                                                          : block_reduce_sum<int64_t, 128>(v2[q], red);
            if (threadIdx.x == 0 && v) atomicAdd(&pc[q], (unsigned long long)v);
        }
    }
}

// jet_refiner.hpp:449-481.
template <bool LDS>
__global__ void __launch_bounds__(SEL_BLOCK_MAX)
    k_strong_select(DGraph g, const part_t* part, const int64_t* sizes, int32_t k, int64_t size_max, int32_t opt,
                    int64_t max_dest, const int64_t* acc, const eoff_t* conn_off, const int2* conn,
                    const int32_t* conn_size, int32_t sections, int32_t* bid, unsigned long long* buckets_base,
                    unsigned long long* extra, int32_t copies, int64_t stride, int32_t* next_counters,
                    BestCopy bc) {
    if (blockIdx.x == 0 && threadIdx.x < 8) next_counters[threadIdx.x] = 0;
    bc.sizes_block0();
    unsigned long long* buckets = bucket_copy(buckets_base, extra, copies, stride);
    __shared__ int32_t bk[BH_SLOTS];
    __shared__ unsigned long long bv[BH_SLOTS];
    __shared__ int64_t ssz[LDS ? SEL_LDS_K : 1];
    PartSizes<LDS> ps{sizes, ssz};
    ps.stage(k);
    BucketMap bm{bk, bv};
    bm.init();
    const int64_t max_vwgt = acc[3];
    FOR_BLOCK_RANGE(i, g.n) {
        const part_t p = part[i];
        if (bc.part_out) bc.part_out[i] = p;
        int32_t b = -1;
        const int64_t sp = ps[p];
        const bool over = sp > size_max;
        const int64_t w = over ? g.vw[i] : 0;
        const eoff_t off = over ? conn_off[i] : 0;
        const int32_t sz = over ? conn_size[i] : 0;
        bool vw_ok = max_vwgt > 0 && w <= 2 * max_vwgt;
        if (over && vw_ok && w < 2 * (sp - opt)) {
            uint64_t tk = 0, tg = 0;
            int64_t p_gain = 0;
            for_slots(conn + off, sz, [&](int2 e) {
                if (e.x == p) {
                    p_gain = e.y;
                } else if (e.x > NULL_PART && ps[e.x] < max_dest) {
                    tg += (uint64_t)e.y;
                    tk += 1;
                }
            });
            if (tk == 0) tk = 1;
            int64_t gain = (int64_t)(tg / tk) - p_gain;
            int64_t vwgt = w < sp - size_max ? w : sp - size_max;
            int32_t gt = gain_bucket_fast(gain, vwgt);
            if (gt < MAX_BUCKETS) {
                b = (MAX_BUCKETS * p + gt) * sections + (int32_t)((uint32_t)i % (uint32_t)sections) + 1;
                bm.add(b, (unsigned long long)w, buckets);
            }
        }
        bid[i] = b;
    }
    bm.flush(buckets);
}

// off = exclusive scan of the bucket sums. WEAK: evicts whole buckets below the
// limit and sends the members of straddling ones to the ranking list; STRONG:
// sends the members of every bucket below the limit there. All members of a
// bucket share its part and offsets, so each decides for itself. One vertex
// per lane: a lane's loads form one dependent chain (measured 0.4 ms less
// per 128^3 stencil partition than four vertices per lane).
// FUSED (nbk <= FUSE_SCAN_MAX): `off` holds the bucket sums, not their scan;
// every workgroup scans them in LDS (one load round trip instead of a scan
// launch) and workgroup 0 writes the scan to off_out for the ranking and
// destination kernels.
// Whether weak bucket b straddles the eviction limit of a part it holds
// members of: its owner q = b / width (gain buckets 0..49), and part q - 1 when
// b is among q's first `sections` minibuckets (the uncapped gain bucket 50 of
// part q - 1 lands there, jet_refiner.hpp:349-351; for q = k, past the last
// part, only part k - 1 holds it).
__device__ inline bool weak_straddle_of(int32_t b, int32_t q, const int64_t* off, const int64_t* sizes,
                                        int64_t size_max, int64_t width) {
    const int64_t base = off[b] - off[(int64_t)q * width];
    const int64_t limit = sizes[q] - size_max;
    return base < limit && base + (off[b + 1] - off[b]) > limit;
}
__device__ inline bool weak_bucket_straddles(int32_t b, int32_t k, int32_t sections, const int64_t* off,
                                             const int64_t* sizes, int64_t size_max, int64_t width) {
    const int32_t q = (int32_t)(b / width);
    if (q < k && weak_straddle_of(b, q, off, sizes, size_max, width)) return true;
    return q >= 1 && (int64_t)b - (int64_t)q * width < sections &&
           weak_straddle_of(b, q - 1, off, sizes, size_max, width);
}

constexpr int CL_CHW = WAVE;
constexpr int CL_WAVES = 16;  // fewer workgroups appending to the list counters (0.2 ms less than 4)
constexpr int64_t FUSE_SCAN_MAX = 4096;
template <bool STRONG, bool FUSED>
__global__ void __launch_bounds__(CL_WAVES* WAVE)
    k_classify(DGraph g, const part_t* part, const int32_t* bid, const int64_t* off_in, int64_t nbk,
               int64_t* off_out, const int64_t* sizes, int64_t size_max, int64_t width, int32_t k, int32_t sections,
               vid_t* moves, int32_t* counters, vid_t nmax, vid_t* rlist, int32_t* stamp, int32_t gen) {
    constexpr int BLOCK = CL_WAVES * WAVE;
    __shared__ int64_t soff[FUSED ? FUSE_SCAN_MAX : 1];
    __shared__ int64_t sws[BLOCK / WAVE + 1];
    const int64_t* off = off_in;
    if (FUSED) {
        constexpr int PER = (int)(FUSE_SCAN_MAX / BLOCK);
        int64_t loc[PER];
        int64_t sum = 0;
        const int64_t t0 = (int64_t)threadIdx.x * PER;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int64_t v = t0 + j < nbk ? off_in[t0 + j] : 0;
            loc[j] = sum;
            sum += v;
        }
        int64_t tot;
        const int64_t ex = block_excl_scan<int64_t, BLOCK>(sum, tot, sws);
#pragma unroll
        for (int j = 0; j < PER; j++)
            if (t0 + j < nbk) soff[t0 + j] = ex + loc[j];
        __syncthreads();
        if (blockIdx.x == 0)
            for (int64_t t = threadIdx.x; t < nbk; t += BLOCK) off_out[t] = soff[t];
        off = soff;
    }
    __shared__ vid_t sbuf[CL_WAVES * CL_CHW];
    __shared__ vid_t rbuf[CL_WAVES * CL_CHW];
    __shared__ int32_t sh[2 * CL_WAVES + 2];
    __shared__ int32_t rsh[2 * CL_WAVES + 2];
    StagedLists<CL_WAVES, CL_CHW> st(sbuf);
    StagedLists<CL_WAVES, CL_CHW> rt(rbuf);
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t wbase = ((int64_t)blockIdx.x * CL_WAVES + threadIdx.x / WAVE) * CL_CHW;
    for (int32_t o = 0; o < CL_CHW; o += WAVE) {
        int64_t i = wbase + o + lane;
        bool ev = false, rank = false;
        eoff_t deg = 0;
        if (i < g.n) {
            int32_t b = bid[i];
            if (b >= 0) {
                part_t p = part[i];
                int64_t base = off[b] - off[(int64_t)p * width];
                int64_t limit = sizes[p] - size_max;
                int64_t bs = off[b + 1] - off[b];
                if (STRONG) {
                    rank = base < limit;
                } else if (base + bs <= limit) {
                    ev = true;
                    deg = g.rowptr[i + 1] - g.rowptr[i];
                }
                // a weak bucket that straddles its part's limit is ranked
                // with all of its members: bucket 50 of part q shares its
                // index with part q + 1's first gain bucket, and a member's
                // position counts every member of lower id (:363-384)
                if (!STRONG) rank = weak_bucket_straddles(b, k, sections, off, sizes, size_max, width);
            }
        }
        st.push(ev && deg <= BIG_DEG, ev && deg > BIG_DEG, (vid_t)i);
        rt.push(rank, false, (vid_t)i);
    }
    st.flush_stamped(moves, &counters[2], &counters[3], nmax, sh, stamp, gen);
    rt.flush(rlist, &counters[5], &counters[7], g.n, rsh);
}

// Decision for one ranked vertex: pos = weight of same-bucket members with a
// smaller id (of any part). Returns 0 (kept), 1 (evicted, degree <= BIG_DEG) or 2 (evicted,
// larger); the caller appends evicted members to the move list.
template <bool STRONG>
__device__ inline int rank_decide_with(vid_t v, int32_t b, int64_t pos, part_t p, int64_t w, bool big,
                                       const int64_t* off, const int64_t* sizes, int64_t size_max, int64_t width,
                                       int64_t* evict_adjust, int64_t* score, int32_t* stamp, int32_t gen) {
    int64_t base = off[b] - off[(int64_t)p * width];
    int64_t limit = sizes[p] - size_max;
    int64_t bs = off[b + 1] - off[b];
    bool ev = false;
    if (STRONG) {
        if (base < limit && base + pos < limit) {
            ev = true;
            if (base + pos + w >= limit) evict_adjust[p] = base + pos + w;
            score[v] = off[b] + pos;
        }
    } else if (base < limit && base + bs > limit) {
        ev = base + pos < limit;
    }
    if (!ev) return 0;
    stamp[v] = gen;
    return big ? 2 : 1;
}

template <bool STRONG>
__device__ inline int rank_decide(vid_t v, int32_t b, int64_t pos, const part_t* part, const int64_t* off,
                                  const int64_t* sizes, int64_t size_max, int64_t width, const wgt_t* vw,
                                  const eoff_t* rowptr, int64_t* evict_adjust, int64_t* score, int32_t* stamp,
                                  int32_t gen) {
    // the member's part, weight and row are loaded together
    const part_t p = part[v];
    const int64_t w = vw[v];
    const eoff_t deg = rowptr[v + 1] - rowptr[v];
    return rank_decide_with<STRONG>(v, b, pos, p, w, deg > BIG_DEG, off, sizes, size_max, width, evict_adjust, score,
                                    stamp, gen);
}


// The cookie cutter (jet_refiner.hpp:491-518) by one workgroup of BLOCK
// threads: two scans over the parts give each part's eviction adjustment and
// the destination chunk starts; the re-draw counter starts after every
// eviction.
template <int BLOCK>
__device__ inline void cookie_block(int32_t k, int64_t width, const int64_t* off, const int64_t* sizes,
                                    int64_t max_dest, int64_t* evict_adjust, int64_t* evict_start, int64_t* acc,
                                    int64_t* ws) {
    int64_t run = 0, run2 = 0;
    for (int32_t p0 = 0; p0 < k; p0 += BLOCK) {
        int32_t p = p0 + threadIdx.x;
        int64_t add = 0, cap = 0, beg = 0;
        if (p < k) {
            int64_t begin = (int64_t)p * width;
            beg = off[begin];
            add = evict_adjust[p];
            if (add == 0) add = off[begin + width] - beg;
            if (max_dest > sizes[p]) cap = max_dest - sizes[p];
        }
        int64_t t1, t2;
        int64_t e1 = block_excl_scan<int64_t, BLOCK>(add, t1, ws);
        int64_t e2 = block_excl_scan<int64_t, BLOCK>(cap, t2, ws);
        if (p < k) {
            evict_adjust[p] = beg - (run + e1);
            evict_start[p + 1] = run2 + e2 + cap;
        }
        run += t1;
        run2 += t2;
    }
    if (threadIdx.x == 0) {
        acc[3] = run;
        evict_start[0] = 0;
    }
}

// What a strong ranking workgroup also does when it ranks every member
// itself: the cookie cutter (so the step needs no cookie launch).
struct CookieArgs {
    int32_t k;  // 0: not here
    int64_t max_dest;
    int64_t* evict_start;
    int64_t* acc;
};

constexpr int RS_CAP = 8192;
constexpr int RS_BLOCK = 1024;

// Ranks up to RS_CAP members in one workgroup: bitonic sort of
// (bucket << 32 | id) in LDS, prefix of weights, decisions.
template <bool STRONG>
__global__ void __launch_bounds__(RS_BLOCK)
    k_rank_small(const vid_t* rlist, const int32_t* counters, const int32_t* bid, const part_t* part,
                 const int64_t* off, const int64_t* sizes, int64_t size_max, int64_t width, const wgt_t* vw,
                 const eoff_t* rowptr, int64_t* evict_adjust, int64_t* score, vid_t* moves, int32_t* mcounters,
                 vid_t nmax, int32_t* deferred, int32_t* stamp, int32_t gen, CookieArgs ck) {
    __shared__ uint64_t keys[RS_CAP];
    __shared__ int64_t pre[RS_CAP];
    __shared__ int64_t ws[RS_BLOCK / WAVE + 1];
    __shared__ uint32_t pay_w[RS_BLOCK];
    __shared__ uint32_t pay_p[RS_BLOCK];
    // each thread's first member is loaded together with the count
    const vid_t m0 = (vid_t)threadIdx.x < nmax ? rlist[threadIdx.x] : 0;
    const int32_t R = counters[5];
    if (R > RS_CAP) {
        // too many members for one workgroup: the host ranks them with the
        // radix path after the iteration's sync; until then the iteration's
        // remaining kernels see the flag and do nothing
        if (threadIdx.x == 0) *deferred = 1;
        return;
    }
    uint32_t cap = 1;
    while ((int32_t)cap < R) cap <<= 1;
    if (cap <= RS_BLOCK) {
        // one member per thread, its weight, part and degree class loaded
        // with its bucket and carried through the sort: exchanges within a
        // wave are shuffles, only strides of a wave or more go through LDS
        // (10 barriers instead of 55 for 1024 keys)
        const uint32_t t = threadIdx.x;
        uint64_t key = ~0ull;
        uint32_t pw = 0, pp = 0;  // weight; part | big << 31
        if ((int32_t)t < R) {
            const int32_t bb = bid[m0];
            const wgt_t w0 = vw[m0];
            const part_t p0 = part[m0];
            const eoff_t d0 = rowptr[m0 + 1] - rowptr[m0];
            key = ((uint64_t)(uint32_t)bb << 32) | (uint32_t)m0;
            pw = (uint32_t)w0;
            pp = (uint32_t)p0 | (d0 > BIG_DEG ? 0x80000000u : 0u);
        }
        for (uint32_t size = 2; size <= cap; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                uint64_t other;
                uint32_t opw, opp;
                if (stride >= WAVE) {
                    keys[t] = key;
                    pay_w[t] = pw;
                    pay_p[t] = pp;
                    __syncthreads();
                    other = keys[t ^ stride];
                    opw = pay_w[t ^ stride];
                    opp = pay_p[t ^ stride];
                    __syncthreads();
                } else {
                    const uint32_t lo = __shfl_xor((uint32_t)key, (int)stride, WAVE);
                    const uint32_t hi = __shfl_xor((uint32_t)(key >> 32), (int)stride, WAVE);
                    other = ((uint64_t)hi << 32) | lo;
                    opw = __shfl_xor(pw, (int)stride, WAVE);
                    opp = __shfl_xor(pp, (int)stride, WAVE);
                }
                // ascending blocks keep the smaller key at the lower index
                const bool lower = (t & stride) == 0, up = (t & size) == 0;
                const bool take = (lower == up) ? (other < key) : (other > key);
                if (take) {
                    key = other;
                    pw = opw;
                    pp = opp;
                }
            }
        }
        keys[t] = key;
        int64_t tot;
        const int64_t ex = block_excl_scan<int64_t, RS_BLOCK>((int32_t)t < R ? (int64_t)(int32_t)pw : 0, tot, ws);
        pre[t] = ex;
        __syncthreads();
        if ((int32_t)t < R) {
            const int32_t b = (int32_t)(key >> 32);
            const uint64_t lo_key = (uint64_t)(uint32_t)b << 32;
            int32_t lo = 0, hi = (int32_t)t;
            while (lo < hi) {
                int32_t mid = (lo + hi) >> 1;
                if (keys[mid] < lo_key) lo = mid + 1;
                else hi = mid;
            }
            const vid_t v = (vid_t)(key & 0xffffffffull);
            const int code = rank_decide_with<STRONG>(v, b, pre[t] - pre[lo], (part_t)(pp & 0x7fffffffu),
                                                      (int64_t)(int32_t)pw, (pp >> 31) != 0, off, sizes, size_max,
                                                      width, evict_adjust, score, stamp, gen);
            if (code == 2) moves[nmax - 1 - atomicAdd(&mcounters[3], 1)] = v;
            else if (code == 1) moves[atomicAdd(&mcounters[2], 1)] = v;
        }
        if (STRONG && ck.k > 0) {
            __syncthreads();  // every member's eviction adjustment written
            cookie_block<RS_BLOCK>(ck.k, width, off, sizes, ck.max_dest, evict_adjust, ck.evict_start, ck.acc, ws);
        }
        return;
    } else {
        for (uint32_t t = threadIdx.x; t < cap; t += RS_BLOCK) {
            if ((int32_t)t < R) {
                const vid_t v = t == threadIdx.x ? m0 : rlist[t];
                keys[t] = ((uint64_t)(uint32_t)bid[v] << 32) | (uint32_t)v;
            } else {
                keys[t] = ~0ull;
            }
        }
        for (uint32_t size = 2; size <= cap; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (uint32_t t = threadIdx.x; t < cap / 2; t += RS_BLOCK) {
                    uint32_t lo = 2 * t - (t & (stride - 1));
                    uint32_t hi = lo + stride;
                    bool up = (lo & size) == 0;
                    uint64_t x = keys[lo], y = keys[hi];
                    if ((x > y) == up) {
                        keys[lo] = y;
                        keys[hi] = x;
                    }
                }
            }
        }
    }
    __syncthreads();
    // exclusive prefix of weights in sorted order (RS_CAP / RS_BLOCK items per thread)
    constexpr int PER = RS_CAP / RS_BLOCK;
    int64_t loc[PER];
    int64_t sum = 0;
    for (int j = 0; j < PER; j++) {
        int32_t t = threadIdx.x * PER + j;
        int64_t w = t < R ? (int64_t)vw[(vid_t)(keys[t] & 0xffffffffull)] : 0;
        loc[j] = sum;
        sum += w;
    }
    int64_t tot;
    int64_t ex = block_excl_scan<int64_t, RS_BLOCK>(sum, tot, ws);
    for (int j = 0; j < PER; j++) pre[threadIdx.x * PER + j] = ex + loc[j];
    __syncthreads();
    for (int32_t t = threadIdx.x; t < R; t += RS_BLOCK) {
        const uint64_t key = keys[t];
        const int32_t b = (int32_t)(key >> 32);
        const uint64_t lo_key = (uint64_t)(uint32_t)b << 32;
        int32_t lo = 0, hi = t;
        while (lo < hi) {
            int32_t mid = (lo + hi) >> 1;
            if (keys[mid] < lo_key) lo = mid + 1;
            else hi = mid;
        }
        const vid_t v = (vid_t)(key & 0xffffffffull);
        const int code = rank_decide<STRONG>(v, b, pre[t] - pre[lo], part, off, sizes, size_max, width, vw, rowptr,
                                             evict_adjust, score, stamp, gen);
        if (code == 2) moves[nmax - 1 - atomicAdd(&mcounters[3], 1)] = v;
        else if (code == 1) moves[atomicAdd(&mcounters[2], 1)] = v;
    }
    if (STRONG && ck.k > 0) {
        __syncthreads();  // every member's eviction adjustment written
        cookie_block<RS_BLOCK>(ck.k, width, off, sizes, ck.max_dest, evict_adjust, ck.evict_start, ck.acc, ws);
    }
}

__global__ void k_rank_keys(const vid_t* rlist, int32_t R, const int32_t* bid, uint64_t* keys, uint32_t* vals) {
    FOR_THREAD(t, R) {
        vid_t v = rlist[t];
        keys[t] = ((uint64_t)(uint32_t)bid[v] << 32) | (uint32_t)v;
        vals[t] = (uint32_t)v;
    }
}

__global__ void k_rank_weights(int32_t R, const uint32_t* vals, const wgt_t* vw, int64_t* w) {
    FOR_THREAD(t, R) w[t] = vw[vals[t]];
}

template <bool STRONG>
__global__ void k_rank_apply(int32_t R, const uint64_t* keys, const int64_t* pre, const part_t* part,
                             const int64_t* off, const int64_t* sizes, int64_t size_max, int64_t width,
                             const wgt_t* vw, const eoff_t* rowptr, int64_t* evict_adjust, int64_t* score,
                             vid_t* moves, int32_t* mcounters, vid_t nmax, int32_t* stamp, int32_t gen) {
    FOR_THREAD(t, R) {
        uint64_t key = keys[t];
        int32_t b = (int32_t)(key >> 32);
        uint64_t lo_key = (uint64_t)(uint32_t)b << 32;
        int64_t lo = 0, hi = t;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < lo_key) lo = mid + 1;
            else hi = mid;
        }
        const vid_t v = (vid_t)(key & 0xffffffffull);
        const int code = rank_decide<STRONG>(v, b, pre[t] - pre[lo], part, off, sizes, size_max, width, vw, rowptr,
                                             evict_adjust, score, stamp, gen);
        if (code == 2) moves[nmax - 1 - atomicAdd(&mcounters[3], 1)] = v;
        else if (code == 1) moves[atomicAdd(&mcounters[2], 1)] = v;
    }
}

// Cookie cutter (jet_refiner.hpp:491-518): per-part eviction totals become a
// global eviction order; receivers get consecutive chunks of capacity. Two
// workgroup scans over the parts.
constexpr int CK_BLOCK = 1024;
__global__ void __launch_bounds__(CK_BLOCK)
    k_cookie(int32_t k, int64_t width, const int64_t* off, const int64_t* sizes, int64_t max_dest,
             int64_t* evict_adjust, int64_t* evict_start, int64_t* acc, const int32_t* deferred) {
    if (*deferred) return;
    __shared__ int64_t ws[CK_BLOCK / WAVE + 1];
    cookie_block<CK_BLOCK>(k, width, off, sizes, max_dest, evict_adjust, evict_start, acc, ws);
}

// Last index q in [lo, k] with evict_start[q] <= s (evict_start is non-decreasing).
__device__ inline int32_t chunk_of(const int64_t* es, int32_t lo, int32_t k, int64_t s) {
    int32_t a = lo, b = k + 1;  // first index with es > s lies in [a, b]
    while (a < b) {
        int32_t mid = (a + b) >> 1;
        if (es[mid] <= s) a = mid + 1;
        else b = mid;
    }
    return a - 1;
}

// jet_refiner.hpp:519-544 without the re-draws, which are queued.
__global__ void k_strong_dest(const vid_t* moves, const int32_t* mc, vid_t nmax, const part_t* part, const wgt_t* vw,
                              const int64_t* evict_adjust, const int64_t* evict_start, int32_t k, int64_t* score,
                              part_t* dest, vid_t* redraw, int32_t* rcount, int32_t rmax,
                              const int32_t* deferred) {
    if (*deferred) return;
    const int32_t ns = mc[2], total = mc[2] + mc[3];
    FOR_THREAD(x, total) {
        vid_t v = list_at(moves, ns, nmax, x);
        part_t pv = part[v];
        int64_t s = score[v] - evict_adjust[pv];
        score[v] = s;
        int32_t q = chunk_of(evict_start, 0, k, s);
        if (q < k && vw[v] / 2 <= evict_start[q + 1] - s) {
            dest[v] = q;
        } else if (q < k) {
            int32_t r = atomicAdd(rcount, 1);
            if (r < rmax) redraw[r] = v;
            dest[v] = pv;
        } else {
            dest[v] = pv;
        }
    }
}

// Re-draws in increasing vertex id, one after another (the counter starts at
// the total eviction weight). The serial walk reads only LDS: the chunk
// starts (k + 1 <= RD_ES_MAX) and, a batch of 256 at a time, the re-drawn
// vertices' scores and weights are staged there first. Up to lds_max (at
// most RD_MAX) re-drawn vertices are sorted by id in LDS; a longer list is
// put in id order in place
// by marking its vertices in `mark` (n entries, all zero between launches)
// and compacting the marks over all vertex ids, clearing them as it goes.
constexpr int RD_MAX = 8192;
constexpr int RD_ES_MAX = 2049;
constexpr int RD_SCAN = 16;  // marks per thread per compaction pass
__global__ void __launch_bounds__(256) k_strong_redraw(vid_t* redraw, const int32_t* rcount, const part_t* part,
                                                       const wgt_t* vw, const int64_t* evict_start, int32_t k,
                                                       const int64_t* score, part_t* dest, int64_t* acc,
                                                       int32_t* mark, vid_t n, int32_t lds_max,
                                                       const int32_t* deferred) {
    __shared__ vid_t ids[RD_MAX];
    __shared__ int64_t es_lds[RD_ES_MAX];
    __shared__ int64_t bscore[256];
    __shared__ wgt_t bw[256];
    __shared__ int64_t red[256 / WAVE + 1];
    if (*deferred) return;
    const int32_t r = rcount[0];
    if (r == 0) return;
    const bool es_in_lds = k + 1 <= RD_ES_MAX;
    if (es_in_lds)
        for (int32_t t = threadIdx.x; t <= k; t += blockDim.x) es_lds[t] = evict_start[t];
    const int64_t* es = es_in_lds ? es_lds : evict_start;
    const vid_t* idp = ids;
    if (r <= lds_max) {
        uint32_t cap = 1;
        while ((int32_t)cap < r) cap <<= 1;
        for (uint32_t t = threadIdx.x; t < cap; t += blockDim.x) ids[t] = (int32_t)t < r ? redraw[t] : ORD_MAX;
        for (uint32_t size = 2; size <= cap; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (uint32_t t = threadIdx.x; t < cap / 2; t += blockDim.x) {
                    uint32_t lo = 2 * t - (t & (stride - 1));
                    uint32_t hi = lo + stride;
                    bool up = (lo & size) == 0;
                    vid_t a = ids[lo], b = ids[hi];
                    if ((a > b) == up) {
                        ids[lo] = b;
                        ids[hi] = a;
                    }
                }
            }
        }
    } else {
        // the list's vertices marked, then written back in increasing id
        for (int32_t x = threadIdx.x; x < r; x += blockDim.x) mark[redraw[x]] = 1;
        __threadfence_block();
        __syncthreads();
        int64_t out = 0;
        for (int64_t v0 = 0; v0 < n; v0 += 256 * RD_SCAN) {
            const int64_t t0 = v0 + (int64_t)threadIdx.x * RD_SCAN;
            int32_t f[RD_SCAN];
            int64_t c = 0;
#pragma unroll
            for (int j = 0; j < RD_SCAN; j++) {
                f[j] = t0 + j < n ? mark[t0 + j] : 0;
                c += f[j];
            }
            int64_t tot;
            int64_t at = out + block_excl_scan<int64_t, 256>(c, tot, red);
#pragma unroll
            for (int j = 0; j < RD_SCAN; j++)
                if (f[j]) {
                    redraw[at++] = (vid_t)(t0 + j);
                    mark[t0 + j] = 0;
                }
            out += tot;
        }
        __threadfence_block();
        idp = redraw;
    }
    __syncthreads();
    int64_t counter = acc[3];  // carried by thread 0
    for (int32_t base = 0; base < r; base += 256) {
        const int32_t nb = r - base < 256 ? r - base : 256;
        if ((int32_t)threadIdx.x < nb) {
            const vid_t v = idp[base + threadIdx.x];
            bscore[threadIdx.x] = score[v];
            bw[threadIdx.x] = vw[v];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int32_t x = 0; x < nb; x++) {
                const vid_t v = idp[base + x];
                const wgt_t wv = bw[x];
                int64_t sc = bscore[x];
                int32_t p = 0;
                bool done = false;
                while (p < k) {
                    p = chunk_of(es, p, k, sc);
                    if (p < k && wv / 2 <= es[p + 1] - sc) {
                        dest[v] = p;
                        done = true;
                        break;
                    } else if (p < k) {
                        sc = counter;
                        counter += wv;
                    }
                }
                if (!done) dest[v] = part[v];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[3] = counter;
}

__global__ void k_collect_moves(const vid_t* moves, const int32_t* mc, vid_t nmax, const part_t* dest, vid_t* out_v,
                                part_t* out_d) {
    const int32_t total = mc[2] + mc[3];
    FOR_THREAD(x, total) {
        vid_t v = list_at(moves, mc[2], nmax, x);
        out_v[x] = v;
        out_d[x] = dest[v];
    }
}

// ---------------------------------------------------------------- host side
static inline int32_t sections_for(vid_t n, int32_t k) {
    int32_t sections = MAX_SECTIONS;
    int32_t section_size = (n + sections * k) / (sections * k);
    if (section_size < 4096) {
        section_size = 4096;
        sections = (n + section_size * k) / (section_size * k);
    }
    return sections;
}

// Lanes per vertex of a list: the average degree rounded up to a power of
// two.
static int group_for_avg(const DGraph& g) {
    double avg = g.n > 0 ? (double)g.m / g.n : 0.0;
    return avg <= 4 ? 4 : avg <= 8 ? 8 : avg <= 16 ? 16 : avg <= 32 ? 32 : 64;
}

// Lanes per candidate in the afterburner: doubled on levels of up to
// SMALL_LEVEL_N vertices, whose few candidates leave the GPU idle otherwise.
// Lanes per move in the update passes: doubled on every level, so that a
// lane's chain of table updates is shorter (each measured 0.5 ms less per
// 128^3 stencil partition; doubling the afterburner's on large levels was
// slower).
static int group_for_afterburner(const DGraph& g) {
    const int m = g.n <= tuning().small_level_n ? (int)tuning().ab_lanes_small : 1;
    return std::min(64, m * group_for_avg(g));
}
static int group_for_update(const DGraph& g) { return std::min(64, (int)tuning().upd_lanes * group_for_avg(g)); }

__global__ void __launch_bounds__(256) k_max_degree(DGraph g, int32_t* out) {
    __shared__ int32_t red[256 / WAVE];
    int32_t mx = 0;
    FOR_THREAD(i, g.n) {
        int64_t d = g.rowptr[i + 1] - g.rowptr[i];
        if (d > mx) mx = (int32_t)(d > ORD_MAX ? ORD_MAX : d);
    }
    mx = block_reduce_max<int32_t, 256>(mx, red);
    if (threadIdx.x == 0 && mx) atomicMax(out, mx);
}

Refiner::Refiner(Arena& a, hipStream_t s, vid_t n, int64_t cap, const RefineParams& p) : a_(a), s_(s), p_(p), nmax_(n) {
    if (p.k < 1 || p.k > MAX_K) throw std::invalid_argument("k out of range");
    if (2 * sizeof(int2) * (size_t)p.k > 65536) {
        // dynamic LDS above 64 KB must be allowed per kernel (gfx950 has 160 KB)
        JET_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conn_build_block),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)(2 * sizeof(int2) * MAX_K)));
    }
    int64_t N = std::max<int64_t>(n, 1);
    conn_off_ = a.alloc<eoff_t>(N + 1);
    conn_ = a.alloc<int2>(std::max<int64_t>(cap, 1));
    conn_size_ = a.alloc<int32_t>(N);
    dest_cache_ = a.alloc<part_t>(N);
    lock_ = a.alloc<int32_t>(N);
    gain_p_ = a.alloc<int32_t>(N);
    pregain_ = a.alloc<int32_t>(N);
    dest_ = a.alloc<part_t>(N);
    bid_ = a.alloc<int32_t>(N);
    cand_ = a.alloc<vid_t>(N);
    moves_ = a.alloc<vid_t>(N);
    mrec_ = a.alloc<MoveRec>(N);
    swap_bit_ = a.alloc<int32_t>(N);
    sval_ = a.alloc<uint32_t>(N);
    vtmp_ = a.alloc<uint32_t>(N);
    prefix_ = a.alloc<int64_t>(N + 1);
    score_ = a.alloc<int64_t>(N);
    cur_part_ = a.alloc<part_t>(N);
    tier_mid_ = a.alloc<vid_t>(N);
    tier_big_ = a.alloc<vid_t>(N);
    hub_ = a.alloc<uint8_t>(N);
    hub_dirty_ = a.alloc<uint8_t>(N);
    pull_mid_ = a.alloc<vid_t>(N);
    stamp_ = a.alloc<int32_t>(N);
    counters_ = a.alloc<int32_t>(32);
    acc_ = a.alloc<int64_t>(8);
    curr_sizes_ = a.alloc<int64_t>(p.k);
    best_sizes_ = a.alloc<int64_t>(p.k);
    evict_adjust_ = a.alloc<int64_t>(p.k);
    evict_start_ = a.alloc<int64_t>(p.k + 1);
    undersized_ = a.alloc<part_t>(p.k);
    misc_ = a.alloc<int32_t>(8);
    redraw_ = a.alloc<vid_t>(N);
    rd_mark_ = a.alloc<int32_t>(N);
    nbuckets_ = (int64_t)MAX_BUCKETS * p.k * sections_for(n, p.k) + MAX_SECTIONS + 4;
    buckets_ = a.alloc<int64_t>(nbuckets_);
    boff_ = a.alloc<int64_t>(std::max<int64_t>(FUSE_SCAN_MAX, nbuckets_));
    ncopies_ = (int32_t)std::max<int64_t>(
        1, std::min<int64_t>(BUCKET_COPIES_MAX, BUCKET_COPIES_BYTES / ((int64_t)sizeof(int64_t) * nbuckets_)));
    bcopies_ = a.alloc<int64_t>(std::max<int64_t>(1, nbuckets_ * (ncopies_ - 1)));
    if (ncopies_ > 1) fill_i64(bcopies_, 0, nbuckets_ * (ncopies_ - 1), s);
    fill_i32(swap_bit_, 0, N, s);
    fill_i32(rd_mark_, 0, N, s);
    JET_HIP(hipMemsetAsync(hub_dirty_, 0, (size_t)N, s));
    fill_i32(stamp_, -1, N, s);
    fill_i32(counters_, 0, 32, s);
    fill_i64(acc_, 0, 8, s);
    wg_state_ = a.alloc<WgState>(1);
    wg_list_ = a.alloc<vid_t>(N);
    wg_runb_ = a.alloc<int32_t>(WG_MAX_NBK + 1);
    pub_ = a.alloc<int64_t>((int64_t)PUB_GROUPS * PUB_STRIDE);
    fill_i64(pub_, 0, (int64_t)PUB_GROUPS * PUB_STRIDE, s);
    szc_ = a.alloc<int64_t>((int64_t)SIZE_COPIES_MAX * p.k);
    fill_i64(szc_, 0, (int64_t)SIZE_COPIES_MAX * p.k, s);
    // (cut change, largest part) of each iteration land in mailbox slots 8, 9
    Mailbox& mb = thread_mailbox();
    mailbox_ = &mb;
    res_host_ = mb.host + 8;
    res_dev_ = mb.dev + 8;
}

void Refiner::build_tiers(const Level& L) {
    DGraph g = L.view();
    fill_i32(counters_ + 16, 0, 4, s_);
    // on small levels every row with edges takes the wave tier: a thread per
    // row walks its edges one after another, a wave inserts them all at once
    small_max_ = g.n <= tuning().small_level_n ? 0 : SMALL_DEG;
    k_conn_tiers<<<grid_for(g.n, 256, 8192), 256, 0, s_>>>(g, p_.k, tier_mid_, tier_big_, pull_mid_,
                                                            counters_ + 16, hub_, small_max_);
    k_max_degree<<<grid_for(g.n, 256, 512), 256, 0, s_>>>(g, counters_ + 18);
    JET_LAUNCH_CHECK();
    int32_t c[4];
    read_i32_words(counters_ + 16, 4, c, s_);
    ntier_mid_ = c[0];
    ntier_big_ = c[1];
    has_big_ = c[2] > BIG_DEG;
    npull_mid_ = c[3];
}

void Refiner::init_conn(const Level& L, const part_t* part, const int32_t* mark) {
    DGraph g = L.view();
    int32_t* mk = const_cast<int32_t*>(mark);
    if (!mark) {
        ArenaScope sc(a_);
        int32_t* caps = a_.alloc<int32_t>((int64_t)g.n + 1);
        k_conn_caps<<<grid_for(g.n, 256, 8192), 256, 0, s_>>>(g, p_.k, caps);
        JET_LAUNCH_CHECK();
        excl_scan_i32_i64(caps, conn_off_, g.n, a_, s_);
        fill2_i32(dest_cache_, NULL_PART, g.n, lock_, 0, g.n, s_);
    }
    k_conn_build_small<<<grid_for(g.n, SB_BLOCK, 8192), SB_BLOCK, 0, s_>>>(g, part, p_.k, conn_off_, conn_,
                                                                          conn_size_, mk, dest_cache_, nullptr,
                                                                          small_max_);
    JET_LAUNCH_CHECK();
    if (ntier_mid_ > 0) {
        const size_t lds = 2 * sizeof(int2) * (size_t)WB_WAVES * std::min(p_.k, WAVE_TABLE_MAX);
        k_conn_build_wave<<<grid_for(ntier_mid_, WB_WAVES, 8192), WB_WAVES * WAVE, lds, s_>>>(
            g, part, p_.k, conn_off_, conn_, conn_size_, mk, dest_cache_, tier_mid_, ntier_mid_, nullptr);
        JET_LAUNCH_CHECK();
    }
    if (ntier_big_ > 0) {
        // a row's capacity is min(deg, k) <= k: two tables of k slots suffice
        const size_t lds = 2 * sizeof(int2) * (size_t)p_.k;
        k_conn_build_block<<<std::min(ntier_big_, 2048), BB_BLOCK, lds, s_>>>(
            g, part, p_.k, conn_off_, conn_, conn_size_, mk, dest_cache_, tier_big_, ntier_big_, nullptr);
        JET_LAUNCH_CHECK();
    }
}

// The best-state copy due from the previous step, for the next selection
// kernel to make (none when not due); the due flag is cleared.
BestCopy Refiner::take_best_copy() {
    BestCopy b{nullptr, nullptr, nullptr, 0};
    if (best_due_) b = BestCopy{best_out_, best_sizes_, curr_sizes_, p_.k};
    best_due_ = false;
    return b;
}

void Refiner::jet_lp(const Level& L, const part_t* part, double fr) {
    DGraph g = L.view();
    int32_t* c = cnt();
    gen_++;  // this step's moves are stamped with a new generation
    const BestCopy bcp = take_best_copy();
    // a pass per wave on levels that fill the GPU with single-pass waves, two
    // passes per wave in 16-wave workgroups above that (see LS_MULTI_CHW)
    if (g.n <= tuning().lp_one_pass_n) {
        int nb = (int)std::max<int64_t>(1, ceil_div(g.n, (int64_t)LS_WAVES * WAVE));
        family_launch(PF_LP_SELECT, k_lp_select<WAVE, LS_WAVES>, dim3(nb), dim3(LS_WAVES * WAVE), 0, s_, g, part,
                      conn_off_, conn_, conn_size_, dest_cache_, gain_p_, lock_, fr, dest_, pregain_, cand_, c, nmax_,
                      cnt_next(), prof_counters(PF_LP_SELECT), bcp);
    } else {
        int nb = (int)std::max<int64_t>(1, ceil_div(g.n, (int64_t)LS_MULTI_WAVES * LS_MULTI_CHW));
        family_launch(PF_LP_SELECT, k_lp_select<LS_MULTI_CHW, LS_MULTI_WAVES>, dim3(nb), dim3(LS_MULTI_WAVES * WAVE),
                      0, s_, g, part, conn_off_, conn_, conn_size_, dest_cache_, gain_p_, lock_, fr, dest_, pregain_,
                      cand_, c, nmax_, cnt_next(), prof_counters(PF_LP_SELECT), bcp);
    }
    JET_LAUNCH_CHECK();
    // per vertex: cached destination 4, lock 4 (and part 4, best copy 4 when
    // the previous step's best state rides along)
    prof_bytes(PF_LP_SELECT, (bcp.part_out ? 16.0 : 8.0) * g.n);
    if (unsigned long long* pc = prof_counters(PF_AFTERBURNER)) {
        // counting runs only: candidate degrees (the list is complete here)
        k_prof_degsum<<<grid_for(g.n, 256, 1024), 256, 0, s_>>>(g, cand_, c, 0, 1, nmax_, nullptr, pc);
    }
    unsigned ab = (unsigned)std::max<int64_t>(256, ceil_div((int64_t)g.n, (int64_t)AB_WAVES * AB_CHW));
#define JET_AB(GG)                                                                                                    \
    family_launch(PF_AFTERBURNER, k_afterburner_small<GG>, dim3(ab), dim3(AB_WAVES * WAVE), 0, s_, g, cand_, c, part, \
                  dest_, pregain_, lock_, moves_, c, nmax_, stamp_, gen_)
    switch (group_for_afterburner(g)) {
        case 4: JET_AB(4); break;
        case 8: JET_AB(8); break;
        case 16: JET_AB(16); break;
        case 32: JET_AB(32); break;
        default: JET_AB(64); break;
    }
#undef JET_AB
    if (has_big_)
        family_launch(PF_AFTERBURNER, k_afterburner_big, dim3(grid_for((int64_t)g.n * WAVE, 256, 2048)), dim3(256), 0,
                      s_, g, cand_, c, part, dest_, pregain_, lock_, moves_, c, nmax_, stamp_, gen_);
    JET_LAUNCH_CHECK();
}

// Ranks the members of flagged buckets (counter 5 of them) and applies the
// eviction decisions. The single-workgroup ranker is launched without reading
// the count on the host; when the count exceeds its capacity it raises the
// deferral flag (misc_[2]), the iteration's remaining kernels do nothing, and
// rank_deferred() finishes the step with the radix path after the sync.
void Refiner::rank_flagged(const Level& L, const part_t* part, const int64_t* sizes, int64_t width, bool strong) {
    DGraph g = L.view();
    int32_t* c = cnt();
    int64_t smax = size_max_;
    if (strong) {
        // the ranking workgroup also runs the cookie cutter (unless it defers)
        const CookieArgs ck{p_.k, rebalance_shape(g.n, true).max_dest, evict_start_, acc_};
        k_rank_small<true><<<1, RS_BLOCK, 0, s_>>>(cand_, c, bid_, part, off_, sizes, smax, width, g.vw, g.rowptr,
                                                  evict_adjust_, score_, moves_, c, nmax_, misc_ + 2, stamp_, gen_,
                                                  ck);
    } else {
        k_rank_small<false><<<1, RS_BLOCK, 0, s_>>>(cand_, c, bid_, part, off_, sizes, smax, width, g.vw,
                                                   g.rowptr, evict_adjust_, score_, moves_, c, nmax_, misc_ + 2, stamp_,
                                                   gen_, CookieArgs{});
    }
    JET_LAUNCH_CHECK();
}

void Refiner::rank_radix(const Level& L, const part_t* part, const int64_t* sizes, int64_t width, int key_bits,
                         bool strong) {
    DGraph g = L.view();
    int32_t* c = cnt();
    int64_t smax = size_max_;
    int32_t R = read_scalar(c + 5, s_);
    if (R == 0) return;
    ArenaScope sc(a_);
    uint64_t* keys = a_.alloc<uint64_t>(R);
    uint64_t* ktmp = a_.alloc<uint64_t>(R);
    int rb = grid_for(R, 256, 8192);
    k_rank_keys<<<rb, 256, 0, s_>>>(cand_, R, bid_, keys, sval_);
    JET_LAUNCH_CHECK();
    radix_sort_u64(keys, sval_, R, 32 + key_bits, ktmp, vtmp_, a_, s_);
    k_rank_weights<<<rb, 256, 0, s_>>>(R, sval_, g.vw, prefix_);
    JET_LAUNCH_CHECK();
    excl_scan_i64_i64(prefix_, prefix_, R, a_, s_);
    if (strong)
        k_rank_apply<true><<<rb, 256, 0, s_>>>(R, keys, prefix_, part, off_, sizes, smax, width, g.vw, g.rowptr,
                                              evict_adjust_, score_, moves_, c, nmax_, stamp_, gen_);
    else
        k_rank_apply<false><<<rb, 256, 0, s_>>>(R, keys, prefix_, part, off_, sizes, smax, width, g.vw, g.rowptr,
                                               evict_adjust_, score_, moves_, c, nmax_, stamp_, gen_);
    JET_LAUNCH_CHECK();
}

Refiner::RebalanceShape Refiner::rebalance_shape(vid_t n, bool strong) const {
    RebalanceShape r;
    const int32_t k = p_.k;
    r.sections = sections_for(n, k);
    r.width = (int64_t)MAX_BUCKETS * r.sections;
    // weak: gain bucket 50 is not capped (jet_refiner.hpp:349-351) and spills
    // past the last part's minibuckets, so `sections` more are kept and scanned
    r.nbk = strong ? (int64_t)k * r.width + 1 : (int64_t)k * r.width + r.sections + 1;
    const int64_t smax = size_max_;
    if (strong) {
        r.max_dest = std::max<int64_t>((int64_t)opt_ + 1, (int64_t)(smax * 0.99));
    } else {
        r.max_dest = (int64_t)(smax * 0.99);
        if (r.max_dest < smax - 100) r.max_dest = smax - 100;
    }
    r.key_bits = 1;
    while ((int64_t(1) << r.key_bits) <= r.nbk) r.key_bits++;
    return r;
}

// rebalance_strong's destination step (jet_refiner.hpp:480-544), after ranking.
void Refiner::strong_tail(const Level& L, const part_t* part, const int64_t* sizes, const RebalanceShape& r,
                          bool cookie_done) {
    DGraph g = L.view();
    int32_t* c = cnt();
    const int32_t k = p_.k;
    int nb = grid_for(g.n, 256, 8192);
    if (!cookie_done)
        k_cookie<<<1, CK_BLOCK, 0, s_>>>(k, r.width, off_, sizes, r.max_dest, evict_adjust_, evict_start_, acc_,
                                         misc_ + 2);
    k_strong_dest<<<nb, 256, 0, s_>>>(moves_, c, nmax_, part, g.vw, evict_adjust_, evict_start_, k, score_, dest_,
                                      redraw_, c + 4, nmax_, misc_ + 2);
    const int32_t lds_max = (int32_t)std::min<int64_t>(RD_MAX, tuning().redraw_lds_max);
    k_strong_redraw<<<1, 256, 0, s_>>>(redraw_, c + 4, part, g.vw, evict_start_, k, score_, dest_, acc_, rd_mark_,
                                       g.n, lds_max, misc_ + 2);
    JET_LAUNCH_CHECK();
}

// Finishes a step whose ranking was deferred: clears the flag, ranks with the
// radix path and runs the strong destination step.
void Refiner::rank_deferred(const Level& L, const part_t* part, const int64_t* sizes, bool strong) {
    fill_i32(misc_ + 2, 0, 1, s_);
    RebalanceShape r = rebalance_shape(L.n, strong);
    rank_radix(L, part, sizes, r.width, r.key_bits, strong);
    if (strong) strong_tail(L, part, sizes, r, false);
}

void Refiner::rebalance(const Level& L, const part_t* part, const int64_t* sizes, bool strong) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    int32_t* c = cnt();
    const RebalanceShape r = rebalance_shape(g.n, strong);
    const int32_t sections = r.sections;
    const int64_t width = r.width, nbk = r.nbk, max_dest = r.max_dest;
    const int64_t smax = size_max_;
    gen_++;  // this step's moves are stamped with a new generation
    // the previous step's second update pass prepared this kind of step (bucket
    // sums zeroed, undersized parts or capacity), else a reset launch does
    if (!(prep_ready_ && prep_strong_ == strong))
        k_reset_buckets<<<grid_for(nbk + 1, 256, 2048), 256, 0, s_>>>(buckets_, nbk + 1, cnt_next(), sizes, k,
                                                                        max_dest, strong ? 1 : 0, undersized_, misc_,
                                                                        acc_, evict_adjust_);
    prep_ready_ = false;
    const BestCopy bcp = take_best_copy();
    // contiguous ranges of at least 512 vertices per workgroup (128 on small
    // levels, one per thread of a 128-thread workgroup): vertices of one
    // oversized part share few buckets, and the serial per-vertex chain of a
    // heavy range stays short
    int nsel = grid_for(g.n, g.n <= tuning().small_level_n ? 128 : 512, (int)tuning().sel_max_blocks);
    unsigned long long* bk = reinterpret_cast<unsigned long long*>(buckets_);
    unsigned long long* bc = reinterpret_cast<unsigned long long*>(bcopies_);
    // one copy per sel_copy_wgs workgroups (256 by default): fewer
    // workgroups meet on a line, and a copy costs a pass of the fold
    const int32_t copies = std::max(1, std::min<int32_t>(ncopies_, (int32_t)(nsel / tuning().sel_copy_wgs)));
    const bool lds = k <= SEL_LDS_K;
    // 512 threads on large levels (half the serial vertices per thread), 128
    // on small ones (more workgroups over few vertices; 0.3 ms less per 128^3
    // stencil partition than 256)
    const int selb = g.n > tuning().small_level_n ? SEL_BLOCK_MAX : 128;
    if (strong) {
        if (lds)
            k_strong_select<true><<<nsel, selb, 0, s_>>>(g, part, sizes, k, smax, opt_, max_dest, acc_, conn_off_,
                                                         conn_, conn_size_, sections, bid_, bk, bc, copies, nbuckets_,
                                                         cnt_next(), bcp);
        else
            k_strong_select<false><<<nsel, selb, 0, s_>>>(g, part, sizes, k, smax, opt_, max_dest, acc_, conn_off_,
                                                          conn_, conn_size_, sections, bid_, bk, bc, copies, nbuckets_,
                                                          cnt_next(), bcp);
    } else {
        family_launch(PF_WEAK_SELECT, lds ? k_weak_select<true> : k_weak_select<false>, dim3(nsel), dim3(selb), 0,
                      s_, g, part, sizes, k, smax, opt_, max_dest, undersized_, misc_, conn_off_, conn_, conn_size_,
                      sections, dest_, bid_, bk, bc, copies, nbuckets_, prof_counters(PF_WEAK_SELECT), cnt_next(), bcp);
        // per vertex: part 4, destination 4, bucket id 4
        prof_bytes(PF_WEAK_SELECT, (bcp.part_out ? 16.0 : 12.0) * g.n);
    }
    if (copies > 1)
        k_sum_bucket_copies<<<grid_for(nbk + 1, 256, 2048), 256, 0, s_>>>(buckets_, bcopies_, copies, nbk + 1,
                                                                          nbuckets_);
    JET_LAUNCH_CHECK();
    // the scanned buckets go to boff_: a bucket array of up to FUSE_SCAN_MAX
    // sums is scanned by the classification kernel itself, a longer one by a
    // scan launch (out of place, so its workgroups need not wait on each other)
    const bool fused = nbk <= FUSE_SCAN_MAX;
    off_ = boff_;
    if (!fused) excl_scan_i64_i64(buckets_, boff_, nbk, a_, s_);
    int cb = (int)std::max<int64_t>(1, ceil_div(g.n, (int64_t)CL_WAVES * CL_CHW));
#define JET_CLASSIFY(ST, FU)                                                                                 \
    k_classify<ST, FU><<<cb, CL_WAVES * WAVE, 0, s_>>>(g, part, bid_, FU ? buckets_ : boff_, nbk, boff_, sizes,    \
                                                       smax, width, k, sections, moves_, c, nmax_, cand_, stamp_,  \
                                                       gen_)
    if (strong) {
        if (fused) JET_CLASSIFY(true, true);
        else JET_CLASSIFY(true, false);
    } else {
        if (fused) JET_CLASSIFY(false, true);
        else JET_CLASSIFY(false, false);
    }
#undef JET_CLASSIFY
    JET_LAUNCH_CHECK();
    rank_flagged(L, part, sizes, width, strong);
    // the ranking workgroup ran the cookie cutter (when it deferred, every
    // later kernel of the step does nothing and rank_deferred runs it)
    if (strong) strong_tail(L, part, sizes, r, true);
}

// Waits for the step's sequence word in the mapped mailbox: a poll of host
// memory returns within a microsecond or two of the write, where a stream
// synchronisation sleeps and wakes. Every few thousand polls the stream is
// queried, so an error or a stream that finished without the word ends the
// wait.
void Refiner::wait_published(volatile int64_t* res, int64_t seq) {
    for (uint64_t spins = 1;; spins++) {
        if (res[5] == seq) return;
        __builtin_ia32_pause();
        // a wait longer than about half a millisecond (the device is busy with
        // other work, e.g. concurrent jobs) gives the CPU to other threads
        if (spins >= (1u << 14) && spins % 64 == 0) sched_yield();
        if (spins % 4096 == 0) {
            hipError_t e = hipStreamQuery(s_);
            if (e == hipSuccess) {
                if (res[5] == seq) return;
                throw DeviceError("refinement step finished without publishing its results");
            }
            if (e != hipErrorNotReady) {
                JET_HIP(e);
            }
        }
    }
}

bool Refiner::perform_moves(const Level& L, part_t* part, int64_t* sizes, int64_t& cut_change, int64_t& max_size,
                             int next_strong) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    int32_t* c = cnt();
    unsigned long long* usz = reinterpret_cast<unsigned long long*>(sizes);
    int G = group_for_update(g);
    int nbs = grid_for((int64_t)g.n * G, 256, 1024);
    const int64_t seq = ++mailbox_->seq;
    // size-delta copies for the LDS-aggregated deltas (k <= LDS_K): about
    // 4096 entries in all
    const int32_t nsc = k <= LDS_K ? std::max(1, std::min(SIZE_COPIES_MAX, 4096 / k)) : 0;
    const Publish pb{pub_, sizes, k, res_dev_, c, seq, reinterpret_cast<unsigned long long*>(szc_), nsc};
    // the next step's start in case it is a rebalancing step of kind
    // next_strong (the controller knows the kind before this step's result)
    RebalPrep prep{};
    if (next_strong >= 0) {
        const RebalanceShape nr = rebalance_shape(g.n, next_strong != 0);
        prep = RebalPrep{buckets_, nr.nbk + 1, nr.max_dest, next_strong, undersized_, misc_, acc_, evict_adjust_};
    }
    const RebalPrep no_prep{};
    prep_ready_ = false;
    int nbb = grid_for((int64_t)g.n * WAVE, 256, 512);
    for (int add = 0; add < 2; add++) {
        if (unsigned long long* pc = prof_counters(PF_UPDATE)) {
            // counting runs only: move degrees, once per pass (nothing when
            // the step's ranking was deferred: its kernels then do nothing)
            k_prof_degsum<<<grid_for(g.n, 256, 1024), 256, 0, s_>>>(g, moves_, c, 2, 3, nmax_, misc_ + 2, pc);
        }
#define JET_UPD(GG)                                                                                                   \
    if (add) family_launch(PF_UPDATE, k_update_small<true, GG>, dim3(nbs + 1), dim3(256), 0, s_, g, moves_, mrec_, c, \
                           nmax_, part, dest_, conn_off_, conn_, conn_size_, dest_cache_, k, misc_ + 1, swap_bit_,    \
                           hub_, hub_dirty_, misc_ + 2, stamp_, gen_, usz, pub_, pb, prep);                           \
    else family_launch(PF_UPDATE, k_update_small<false, GG>, dim3(nbs), dim3(256), 0, s_, g, moves_, mrec_, c, nmax_, \
                       part, dest_, conn_off_, conn_, conn_size_, dest_cache_, k, misc_ + 1, swap_bit_, hub_,         \
                       hub_dirty_, misc_ + 2, stamp_, gen_, usz, pub_, pb, no_prep);
        switch (G) {
            case 4: JET_UPD(4) break;
            case 8: JET_UPD(8) break;
            case 16: JET_UPD(16) break;
            case 32: JET_UPD(32) break;
            default: JET_UPD(64) break;
        }
#undef JET_UPD
        if (has_big_) {
            const int32_t slices = (int32_t)tuning().upd_big_slices;
            family_launch(PF_UPDATE, add ? k_update_big<true> : k_update_big<false>, dim3(nbb), dim3(256), 0, s_, g,
                          moves_, mrec_, c, nmax_, part, dest_, conn_off_, conn_, conn_size_, dest_cache_, k,
                          misc_ + 1, swap_bit_, hub_, hub_dirty_, misc_ + 2, stamp_, gen_, usz, pub_, slices);
        }
        // pulled rows: big-tier rows by a 256-thread workgroup, rows of
        // degree (BIG_DEG, HEAVY_DEG_REFINE] by a single-wave workgroup
        const uint32_t lds = (uint32_t)(sizeof(int32_t) * (size_t)k);
        if (ntier_big_ > 0) {
            const int hb = std::min(ntier_big_, 2048);
            family_launch(PF_UPDATE, add ? k_update_hubs<true, 256> : k_update_hubs<false, 256>, dim3(hb), dim3(256),
                          lds, s_, g, tier_big_, ntier_big_, c, stamp_, gen_, hub_dirty_, part, dest_, conn_off_,
                          conn_, conn_size_, dest_cache_, k, misc_ + 1, misc_ + 2);
        }
        if (npull_mid_ > 0) {
            const int hb = std::min(npull_mid_, 8192);
            family_launch(PF_UPDATE, add ? k_update_hubs<true, WAVE> : k_update_hubs<false, WAVE>, dim3(hb),
                          dim3(WAVE), lds, s_, g, pull_mid_, npull_mid_, c, stamp_, gen_, hub_dirty_, part, dest_,
                          conn_off_, conn_, conn_size_, dest_cache_, k, misc_ + 1, misc_ + 2);
        }
        JET_LAUNCH_CHECK();
    }
    volatile int64_t* res = res_host_;
    wait_published(res, seq);
    if (res[2]) return true;  // ranking deferred: nothing applied, counters kept
    cut_change = res[0];
    max_size = res[1];
    if (next_strong >= 0) {
        prep_ready_ = true;
        prep_strong_ = next_strong != 0;
    }
    // update_large: the marked rows are rebuilt before the tables are next
    // read (rebuild_pending), so a step that pushed its updates launches no
    // rebuild kernels
    rebuild_pending_ = res[3] != 0;
    par_ ^= 1;
    return false;
}

// The refinement kernels' error bits, as text.
static std::string refine_error_text(int32_t err) {
    std::string m = "refinement invariant violated:";
    if (err & 1) m += " a moved vertex's old part was missing from a neighbour's connectivity row;";
    if (err & 2) m += " a connectivity row ran out of slots;";
    if (err & 8) m += " a strong re-draw walked past its bound;";
    if (err & ~11) m += " error bits " + std::to_string(err) + ";";
    return m;
}

void Refiner::flush_rebuild(const Level& L) {
    if (!rebuild_pending_) return;
    init_conn(L, cur_part_, swap_bit_);
    rebuild_pending_ = false;
}

LevelRecord Refiner::refine(const Level& L, part_t* best_part) {
    auto t0 = std::chrono::steady_clock::now();
    changed_ = false;
    prep_ready_ = false;
    best_due_ = false;
    DGraph g = L.view();
    const int32_t k = p_.k;
    fill2_i32(misc_, 0, 8, counters_, 0, 16, s_);
    par_ = 0;
    if (!init_) {
        cut_ = total_cut2(g, best_part, a_, s_);
        part_sizes(g, best_part, k, best_sizes_, s_);
        std::vector<int64_t> hs(k);
        JET_HIP(hipMemcpyAsync(hs.data(), best_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToHost, s_));
        JET_HIP(hipStreamSynchronize(s_));
        total_size_ = 0;
        for (int64_t v : hs) total_size_ += v;
    }
    opt_ = (int32_t)(((int32_t)total_size_ + k - 1) / k);
    size_max_ = (int32_t)(opt_ * p_.imb);
    if (!init_) {
        init_ = true;
        std::vector<int64_t> hs(k);
        JET_HIP(hipMemcpyAsync(hs.data(), best_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToHost, s_));
        JET_HIP(hipStreamSynchronize(s_));
        int64_t mx = 0;
        for (int64_t v : hs) mx = std::max(mx, v);
        total_imb_ = mx > opt_ ? mx - opt_ : 0;
    }
    int64_t curr_cut = cut_, curr_imb = total_imb_;
    copy_i64(curr_sizes_, best_sizes_, k, s_);
    const int64_t imb_max = (int64_t)size_max_ - opt_;
    copy_i32(cur_part_, best_part, g.n, s_);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    std::vector<double> temps;
    if (p_.ultra) {
        for (double t = 0.85; t > 0; t -= 0.05) temps.push_back(t);
    } else if (L.uniform) {
        temps.push_back(0.25);
    } else {
        temps.push_back(0.75);
    }
    {
        const RebalanceShape wr = rebalance_shape(g.n, false);
        if (!trace && wg_fits(g.n, k, wr.nbk, wg_threshold()) && (int)temps.size() <= WG_MAX_TEMPS) {
            LevelRecord rec = refine_wg(L, best_part, temps);
            rec.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            return rec;
        }
    }
    LevelRecord rec;
    int balance_counter = 0;
    for (double fr : temps) {
        int count = 0;
        while (count++ <= 11) {
            rec.iterations++;
            bool strong = false;
            flush_rebuild(L);
            if (curr_imb <= imb_max) {
                jet_lp(L, cur_part_, fr);
                balance_counter = 0;
                rec.lp_iterations++;
            } else {
                strong = balance_counter >= 2;
                rebalance(L, cur_part_, curr_sizes_, strong);
                balance_counter++;
            }
            int64_t cc = 0, mx = 0;
            // a rebalancing step after this one is strong iff two rebalancing
            // steps precede it (balance_counter as updated for this step)
            const int next_strong = balance_counter >= 2 ? 1 : 0;
            if (perform_moves(L, cur_part_, curr_sizes_, cc, mx, next_strong)) {
                rank_deferred(L, cur_part_, curr_sizes_, strong);
                if (perform_moves(L, cur_part_, curr_sizes_, cc, mx, next_strong))
                    throw DeviceError("rebalancing step deferred twice");
            }
            curr_imb = mx > opt_ ? mx - opt_ : 0;
            curr_cut -= cc;
            int take = 0;  // 1: better balance, 2: better cut
            if (total_imb_ > imb_max && curr_imb < total_imb_) {
                take = 1;
                count = 0;
            } else if (curr_cut < cut_ && (curr_imb <= imb_max || curr_imb <= total_imb_)) {
                if ((double)curr_cut < p_.tol * (double)cut_) count = 0;
                take = 2;
            }
            if (take) {
                changed_ = true;
                cut_ = curr_cut;
                total_imb_ = curr_imb;
                // the best state (partition and part sizes) is copied by the
                // next step's selection kernel, or at the end of the level
                best_due_ = true;
                best_out_ = best_part;
            }
            if (trace) {
                const int64_t kind = balance_counter == 0 ? 0 : (strong ? 2 : 1);
                trace->push_back({kind, -1, curr_cut, curr_imb, take, count});
            }
        }
    }
    prep_ready_ = false;
    if (best_due_) {
        k_copy_best<<<grid_for(g.n, 256, 8192), 256, 0, s_>>>(best_part, cur_part_, g.n, best_sizes_, curr_sizes_, k);
        JET_LAUNCH_CHECK();
        best_due_ = false;
    }
    flush_rebuild(L);
    int32_t err = read_scalar(misc_ + 1, s_);
    if (err) throw DeviceError(refine_error_text(err));
    rec.cut2 = cut_;
    rec.total_imb = total_imb_;
    rec.n = g.n;
    rec.m = g.m;
    rec.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rec;
}

// The level's iterations in the one-workgroup kernel (csrc/refine_wg.hip),
// from the tables init_conn built. The kernel stops early only when a step
// chose update_large: the marked rows are then rebuilt here, by the same
// multi-workgroup build as flush_rebuild, and the kernel continues.
LevelRecord Refiner::refine_wg(const Level& L, part_t* best_part, const std::vector<double>& temps) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    const RebalanceShape weak = rebalance_shape(g.n, false), strong = rebalance_shape(g.n, true);
    WgParams P{};
    P.g = g;
    P.k = k;
    P.big_deg = (int32_t)tuning().wg_big_deg;
    P.lanes = group_for_avg(g);  // the average degree rounded up to a power of two (4-64)
    P.sections = weak.sections;
    P.nbk = (int32_t)weak.nbk;
    P.opt = opt_;
    P.size_max = size_max_;
    P.imb_max = (int64_t)size_max_ - opt_;
    P.max_dest_weak = weak.max_dest;
    P.max_dest_strong = strong.max_dest;
    P.tol = p_.tol;
    P.ntemps = (int32_t)temps.size();
    for (size_t i = 0; i < temps.size(); i++) P.temps[i] = temps[i];
    P.part = cur_part_;
    P.best_part = best_part;
    P.conn_off = conn_off_;
    P.conn = conn_;
    P.conn_size = conn_size_;
    P.dest_cache = dest_cache_;
    P.lock = lock_;
    P.gain_p = gain_p_;
    P.pregain = pregain_;
    P.dest = dest_;
    P.bid = bid_;
    P.cand = cand_;
    P.moves = moves_;
    P.rlist = wg_list_;
    P.score = score_;
    P.mark = swap_bit_;
    P.runb = wg_runb_;
    P.curr_sizes = curr_sizes_;
    P.best_sizes = best_sizes_;
    P.st = wg_state_;
    wg_init_state(wg_state_, cut_, total_imb_, s_);
    WgState h{};
    for (int launches = 0;; launches++) {
        if (launches > 100000) throw DeviceError("one-workgroup refinement made no progress");
        if (WgCollector* col = wg_collector()) {
            col->run(P, s_, h);  // graph-tagged with the levels of other jobs
        } else {
            launch_refine_wg(P, s_);
            JET_HIP(hipMemcpyAsync(&h, wg_state_, sizeof(h), hipMemcpyDeviceToHost, s_));
            JET_HIP(hipStreamSynchronize(s_));
        }
        if (h.err) throw DeviceError(refine_error_text(h.err));
        if (h.status == WG_STALLED) throw DeviceError("one-workgroup refinement exceeded its iteration bound");
        if (h.status != WG_REBUILD) break;
        init_conn(L, cur_part_, swap_bit_);  // update_large's rebuild of the marked rows
    }
    if (h.changed) changed_ = true;
    cut_ = h.best_cut;
    total_imb_ = h.best_imb;
    LevelRecord rec;
    rec.iterations = h.iterations;
    rec.lp_iterations = h.lp_iterations;
    rec.cut2 = cut_;
    rec.total_imb = total_imb_;
    rec.n = g.n;
    rec.m = g.m;
    return rec;
}

void Refiner::single_step(const Level& L, const part_t* part, int kind, double fr, std::vector<int32_t>& moves,
                          std::vector<int32_t>& dest) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    best_due_ = false;
    fill_i32(misc_, 0, 8, s_);
    fill_i32(counters_, 0, 16, s_);
    par_ = 0;
    copy_i32(cur_part_, part, g.n, s_);
    part_sizes(g, cur_part_, k, curr_sizes_, s_);
    std::vector<int64_t> hs(k);
    JET_HIP(hipMemcpyAsync(hs.data(), curr_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    int64_t tot = 0;
    for (int64_t v : hs) tot += v;
    opt_ = (int32_t)(((int32_t)tot + k - 1) / k);
    size_max_ = (int32_t)(opt_ * p_.imb);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    if (kind == 0) jet_lp(L, cur_part_, fr);
    else {
        rebalance(L, cur_part_, curr_sizes_, kind == 2);
        if (read_scalar(misc_ + 2, s_)) rank_deferred(L, cur_part_, curr_sizes_, kind == 2);
    }
    int32_t c4[4];
    JET_HIP(hipMemcpyAsync(c4, cnt(), sizeof(c4), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    int32_t total = c4[2] + c4[3];
    ArenaScope sc(a_);
    vid_t* ov = a_.alloc<vid_t>(std::max(total, 1));
    part_t* od = a_.alloc<part_t>(std::max(total, 1));
    if (total > 0) {
        k_collect_moves<<<grid_for(total, 256, 8192), 256, 0, s_>>>(moves_, cnt(), nmax_, dest_, ov, od);
        JET_LAUNCH_CHECK();
    }
    std::vector<int32_t> v(total), d(total);
    if (total > 0) {
        JET_HIP(hipMemcpyAsync(v.data(), ov, sizeof(int32_t) * total, hipMemcpyDeviceToHost, s_));
        JET_HIP(hipMemcpyAsync(d.data(), od, sizeof(int32_t) * total, hipMemcpyDeviceToHost, s_));
    }
    JET_HIP(hipStreamSynchronize(s_));
    int32_t err = read_scalar(misc_ + 1, s_);
    if (err) throw DeviceError(refine_error_text(err));
    std::vector<int32_t> idx(total);
    for (int32_t t = 0; t < total; t++) idx[t] = t;
    std::sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return v[x] < v[y]; });
    moves.resize(total);
    dest.resize(total);
    for (int32_t t = 0; t < total; t++) {
        moves[t] = v[idx[t]];
        dest[t] = d[idx[t]];
    }
}

// part_stat.hpp:131-144, 186-239 from the connectivity tables: the
// communication volume (per vertex, the parts other than its own that its
// neighbours are in) and the cut weight incident to each part.
__global__ void k_quality(DGraph g, const part_t* part, const eoff_t* conn_off, const int2* conn,
                          const int32_t* conn_size, unsigned long long* comm, unsigned long long* cut_per_part) {
    __shared__ int64_t red[256 / WAVE];
    int64_t c = 0;
    FOR_THREAD(i, g.n) {
        const part_t p = part[i];
        const int2* tab = conn + conn_off[i];
        const int32_t sz = conn_size[i];
        int64_t ext = 0;
        for (int32_t s = 0; s < sz; s++) {
            int2 e = tab[s];
            if (e.x > NULL_PART && e.x != p && e.y > 0) {
                c++;
                ext += e.y;
            }
        }
        if (ext) atomicAdd(&cut_per_part[p], (unsigned long long)ext);
    }
    c = block_reduce_sum<int64_t, 256>(c, red);
    if (threadIdx.x == 0 && c) atomicAdd(comm, (unsigned long long)c);
}

void Refiner::quality(const Level& L, const part_t* part, int64_t& comm_size, std::vector<int64_t>& cut_per_part) {
    DGraph g = L.view();
    copy_i32(cur_part_, part, g.n, s_);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    ArenaScope sc(a_);
    unsigned long long* d = a_.alloc<unsigned long long>(1 + p_.k);
    fill_i64(reinterpret_cast<int64_t*>(d), 0, 1 + p_.k, s_);
    k_quality<<<grid_for(g.n, 256, 4096), 256, 0, s_>>>(g, cur_part_, conn_off_, conn_, conn_size_, d, d + 1);
    JET_LAUNCH_CHECK();
    std::vector<int64_t> h(1 + p_.k);
    JET_HIP(hipMemcpyAsync(h.data(), d, sizeof(int64_t) * h.size(), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    comm_size = h[0];
    cut_per_part.assign(h.begin() + 1, h.end());
}

// part_stat.hpp:116-129 from the connectivity tables: heat[p * k + q] = cut
// weight from vertices of part p to part q (q != p). With k * k <= HEAT_LDS
// each workgroup sums into LDS first.
constexpr int HEAT_LDS = 4096;

template <bool LDS>
__global__ void __launch_bounds__(256)
    k_heatmap(DGraph g, const part_t* part, const eoff_t* conn_off, const int2* conn, const int32_t* conn_size,
              int32_t k, unsigned long long* heat) {
    __shared__ unsigned long long lh[LDS ? HEAT_LDS : 1];
    if (LDS) {
        for (int t = threadIdx.x; t < k * k; t += blockDim.x) lh[t] = 0;
        __syncthreads();
    }
    FOR_BLOCK_RANGE(i, g.n) {
        const part_t p = part[i];
        for_slots(conn + conn_off[i], conn_size[i], [&](int2 e) {
            if (e.x > NULL_PART && e.x != p && e.y > 0) {
                const int64_t at = (int64_t)p * k + e.x;
                if (LDS) atomicAdd(&lh[at], (unsigned long long)e.y);
                else atomicAdd(&heat[at], (unsigned long long)e.y);
            }
        });
    }
    if (LDS) {
        __syncthreads();
        for (int t = threadIdx.x; t < k * k; t += blockDim.x)
            if (lh[t]) atomicAdd(&heat[t], lh[t]);
    }
}

void Refiner::heatmap(const Level& L, const part_t* part, int64_t* heat) {
    DGraph g = L.view();
    copy_i32(cur_part_, part, g.n, s_);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    const int64_t kk = (int64_t)p_.k * p_.k;
    fill_i64(heat, 0, kk, s_);
    unsigned long long* h = reinterpret_cast<unsigned long long*>(heat);
    if (kk <= HEAT_LDS)
        k_heatmap<true><<<grid_for(g.n, 2048, 1024), 256, 0, s_>>>(g, cur_part_, conn_off_, conn_, conn_size_, p_.k, h);
    else
        k_heatmap<false><<<grid_for(g.n, 256, 8192), 256, 0, s_>>>(g, cur_part_, conn_off_, conn_, conn_size_, p_.k,
                                                                   h);
    JET_LAUNCH_CHECK();
}

void Refiner::conn_tables(const Level& L, const part_t* part, std::vector<int64_t>& off,
                          std::vector<int32_t>& slot_part, std::vector<int32_t>& slot_val, std::vector<int32_t>& size) {
    DGraph g = L.view();
    copy_i32(cur_part_, part, g.n, s_);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    off.resize(g.n + 1);
    size.resize(g.n);
    JET_HIP(hipMemcpyAsync(off.data(), conn_off_, sizeof(int64_t) * (g.n + 1), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipMemcpyAsync(size.data(), conn_size_, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    int64_t C = off[g.n];
    std::vector<int2> slots(std::max<int64_t>(C, 1));
    if (C > 0) JET_HIP(hipMemcpyAsync(slots.data(), conn_, sizeof(int2) * C, hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    slot_part.resize(C);
    slot_val.resize(C);
    for (int64_t c = 0; c < C; c++) {
        slot_part[c] = slots[c].x;
        slot_val[c] = slots[c].y;
    }
}

}  // namespace jet
