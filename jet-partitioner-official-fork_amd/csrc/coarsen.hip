// Coarsening mappers for gfx950: HEC heavy-edge selection, claim-round map
// construction, heavy/random matching with hash-gated commit rounds, and the
// 2-hop leaf/twin/relative stages (reference: src/heuristics.hpp).
//
// Every order-dependent step of the reference (CAS claims, matchHash slot
// races) is computed here as the outcome of one fixed legal schedule, in
// parallel and deterministically:
//  * map construction: all first CASes precede all second steps, which run in
//    decreasing u (swap = 1) or increasing u (swap = 0);
//  * matching commit part 2: the smallest active u claiming v wins;
//  * matchHash: candidates with equal digest pair two at a time in increasing
//    vertex id (a stable sort by digest groups them).
#include "coarsen.h"
#include "prims.h"
#include "wave.h"

namespace jet {

static int group_for(const DGraph& g) {
    double avg = g.n > 0 ? (double)g.m / (double)g.n : 0.0;
    if (avg < 8) return 1;
    if (avg < 32) return 8;
    if (avg < 128) return 32;
    return 64;
}

template <int G, class T>
__device__ inline T group_sum(T x) {
#pragma unroll
    for (int d = G / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, G);
    return x;
}

// ---------------------------------------------------------------- heavy list
__global__ void __launch_bounds__(256) k_heavy_list(DGraph g, vid_t* list, int32_t* cnt) {
    __shared__ int32_t sh[256 / WAVE + 1];
    FOR_BLOCK_UNIFORM(i, g.n) {
        const bool heavy = i < g.n && g.rowptr[i + 1] - g.rowptr[i] > HEAVY_DEG;
        int32_t p = block_append<256>(cnt, heavy, sh);
        if (p >= 0) list[p] = (vid_t)i;
    }
}

int32_t build_heavy_list(const DGraph& g, vid_t* list, Arena& a, hipStream_t s) {
    ArenaScope sc(a);
    int32_t* cnt = a.alloc<int32_t>(1);
    fill_i32(cnt, 0, 1, s);
    k_heavy_list<<<grid_for(g.n, 256, 4096), 256, 0, s>>>(g, list, cnt);
    JET_LAUNCH_CHECK();
    return read_scalar(cnt, s);
}

// ---------------------------------------------------------------- HEC
__global__ void k_hec_uniform(DGraph g, uint64_t seed, uint32_t level, vid_t* hn) {
    FOR_THREAD(i, g.n) {
        uint64_t r = rng(seed, level, SITE_HEC, 0, (uint32_t)i);
        eoff_t s0 = g.rowptr[i];
        int64_t deg = g.rowptr[i + 1] - s0;
        hn[i] = deg > 0 ? g.col[s0 + (int64_t)(r % (uint64_t)deg)] : (vid_t)(r % (uint64_t)g.n);
    }
}

__global__ void k_vw_sum(DGraph g, unsigned long long* out) {
    __shared__ int64_t red[256 / WAVE];
    int64_t s = 0;
    FOR_THREAD(i, g.n) s += g.vw[i];
    s = block_reduce_sum<int64_t, 256>(s, red);
    if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)s);
}

// Heavy-neighbour argmax over eligible edges: max weight, then the LAST edge
// among equal weights (heuristics.hpp:224-233 under serial reduction order).
__device__ inline void hec_better(int64_t& bw, int64_t& bj, int64_t w, int64_t j) {
    if (w > bw || (w == bw && j > bj)) {
        bw = w;
        bj = j;
    }
}

template <int G>
__global__ void k_hec_weighted(DGraph g, const unsigned long long* vwsum, vid_t* hn) {
    const int64_t max_allowed = 6 * (int64_t)(*vwsum) / g.n;
    FOR_GROUP(x, g.n) {
        vid_t i = (vid_t)x;
        eoff_t s0 = g.rowptr[i], e0 = g.rowptr[i + 1];
        if (e0 - s0 > HEAVY_DEG) continue;
        if (e0 > s0 && g.vw[i] < max_allowed) {
            int64_t bw = INT64_MIN, bj = -1;
            for (eoff_t j = s0 + sub_; j < e0; j += G) {
                if (g.vw[g.col[j]] < max_allowed) hec_better(bw, bj, g.ew[j], j);
            }
#pragma unroll
            for (int d = G / 2; d >= 1; d >>= 1) {
                int64_t ow = __shfl_xor(bw, d, G), oj = __shfl_xor(bj, d, G);
                hec_better(bw, bj, ow, oj);
            }
            if (sub_ == 0) hn[i] = bj >= 0 ? g.col[bj] : i;
        } else if (sub_ == 0) {
            hn[i] = i;
        }
    }
}

__global__ void __launch_bounds__(256) k_hec_weighted_heavy(DGraph g, const unsigned long long* vwsum,
                                                           const vid_t* heavy, int32_t nheavy, vid_t* hn) {
    __shared__ int64_t sw[256 / WAVE], sj[256 / WAVE];
    const int64_t max_allowed = 6 * (int64_t)(*vwsum) / g.n;
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    for (int32_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        vid_t i = heavy[h];
        eoff_t s0 = g.rowptr[i], e0 = g.rowptr[i + 1];
        if (!(g.vw[i] < max_allowed)) {
            if (threadIdx.x == 0) hn[i] = i;
            continue;
        }
        int64_t bw = INT64_MIN, bj = -1;
        for (eoff_t j = s0 + threadIdx.x; j < e0; j += blockDim.x)
            if (g.vw[g.col[j]] < max_allowed) hec_better(bw, bj, g.ew[j], j);
        for (int d = WAVE / 2; d >= 1; d >>= 1) {
            int64_t ow = __shfl_xor(bw, d, WAVE), oj = __shfl_xor(bj, d, WAVE);
            hec_better(bw, bj, ow, oj);
        }
        if (lane == 0) {
            sw[wid] = bw;
            sj[wid] = bj;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < 256 / WAVE; w++) hec_better(bw, bj, sw[w], sj[w]);
            hn[i] = bj >= 0 ? g.col[bj] : i;
        }
        __syncthreads();
    }
}

void hec_heavy_neighbours(const DGraph& g, bool uniform, uint64_t seed, int32_t level, vid_t* hn, Arena& a,
                          hipStream_t s) {
    if (g.n <= 0) return;
    if (uniform) {
        k_hec_uniform<<<grid_for(g.n, 256, 8192), 256, 0, s>>>(g, seed, (uint32_t)level, hn);
        JET_LAUNCH_CHECK();
        return;
    }
    ArenaScope sc(a);
    unsigned long long* sum = a.alloc<unsigned long long>(1);
    fill_i64(reinterpret_cast<int64_t*>(sum), 0, 1, s);
    k_vw_sum<<<grid_for(g.n, 256, 1024), 256, 0, s>>>(g, sum);
    JET_LAUNCH_CHECK();
    int G = group_for(g);
    int blocks = grid_for((int64_t)g.n * G, 256, 16384);
    switch (G) {
        case 1: k_hec_weighted<1><<<blocks, 256, 0, s>>>(g, sum, hn); break;
        case 8: k_hec_weighted<8><<<blocks, 256, 0, s>>>(g, sum, hn); break;
        case 32: k_hec_weighted<32><<<blocks, 256, 0, s>>>(g, sum, hn); break;
        default: k_hec_weighted<64><<<blocks, 256, 0, s>>>(g, sum, hn); break;
    }
    JET_LAUNCH_CHECK();
    vid_t* heavy = a.alloc<vid_t>(g.n);
    int32_t nh = build_heavy_list(g, heavy, a, s);
    if (nh > 0) {
        k_hec_weighted_heavy<<<std::min(nh, 1024), 256, 0, s>>>(g, sum, heavy, nh, hn);
        JET_LAUNCH_CHECK();
    }
}

// ---------------------------------------------------------------- map construction
__device__ inline bool mc_acting(vid_t u, vid_t h, int swap) { return (((u < h) ? 1 : 0) ^ swap) != 0; }

// Also clears the round's work-list counter (k_mc_next appends to it; the
// host has read the previous round's count before this launch).
__global__ void k_mc_mark(const vid_t* wl, int32_t len, int32_t round, int32_t* stamp, int32_t* claim,
                          int32_t* cnt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *cnt = 0;
    FOR_THREAD(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        stamp[u] = round;
        claim[u] = ORD_MAX;
    }
}

// swap == 0 rounds: the first (smallest) acting u aiming at a non-acting
// unresolved v claims it.
__global__ void k_mc_claim(const vid_t* wl, int32_t len, int32_t round, const vid_t* hn, const int32_t* stamp,
                           int32_t* claim) {
    FOR_THREAD(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        vid_t v = hn[u];
        if (!(u < v)) continue;
        if (stamp[v] == round && !(v < hn[v])) atomicMin(&claim[v], u);
    }
}

__global__ void k_mc_resolve(const vid_t* wl, int32_t len, int32_t round, int swap, const vid_t* hn,
                             const int32_t* stamp, const int32_t* claim, vid_t* vcmap) {
    FOR_THREAD(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        vid_t v = hn[u];
        if (!mc_acting(u, v, swap)) continue;
        if (v == u) {
            vcmap[u] = u;
        } else if (stamp[v] == round) {
            if (mc_acting(v, hn[v], swap)) {
                vcmap[u] = ORD_MAX;  // target still holds ORD_MAX-1: re-queue
            } else {
                vid_t cv = swap ? v : claim[v];
                vcmap[u] = cv;
                vcmap[v] = cv;
            }
        } else {
            vcmap[u] = vcmap[v];  // target aggregated in an earlier round
        }
    }
}

__global__ void __launch_bounds__(256)
    k_mc_next(const vid_t* wl, int32_t len, vid_t n, const vid_t* vcmap, vid_t* next, int32_t* cnt) {
    __shared__ int32_t sh[256 / WAVE + 1];
    FOR_BLOCK_UNIFORM(x, len) {
        vid_t u = 0;
        bool keep = false;
        if (x < len) {
            u = wl ? wl[x] : (vid_t)x;
            keep = vcmap[u] >= n;
        }
        int32_t p = block_append<256>(cnt, keep, sh);
        if (p >= 0) next[p] = u;
    }
}

// The rounds of map construction from a short work list on, in one
// workgroup: the same mark / claim / resolve / next phases as the kernels
// above, separated by workgroup barriers instead of kernel boundaries, until
// the list is empty. The list's order does not matter to any phase.
constexpr int MC_TAIL_BLOCK = 1024;
constexpr int32_t MC_TAIL_MAX = 8192;  // work lists of at most this many go to the tail kernel
__global__ void __launch_bounds__(MC_TAIL_BLOCK)
    k_mc_tail(const vid_t* wl0, int32_t len0, int32_t round0, int swap0, vid_t n, const vid_t* hn, int32_t* stamp,
              int32_t* claim, vid_t* vcmap, vid_t* wl_a, vid_t* wl_b, int32_t* err) {
    __shared__ int32_t cnt;
    if (threadIdx.x == 0) *err = 0;  // only thread 0 writes it
    const vid_t* wl = wl0;
    int32_t len = len0;
    int swap = swap0;
    for (int32_t round = round0; len > 0; round++) {
        if (round > 2 * n + 8) {  // each round resolves at least one vertex
            if (threadIdx.x == 0) atomicOr(err, 1);
            return;
        }
        for (int32_t x = threadIdx.x; x < len; x += MC_TAIL_BLOCK) {
            const vid_t u = wl[x];
            stamp[u] = round;
            claim[u] = ORD_MAX;
        }
        if (threadIdx.x == 0) cnt = 0;
        __syncthreads();
        if (swap == 0) {
            for (int32_t x = threadIdx.x; x < len; x += MC_TAIL_BLOCK) {
                const vid_t u = wl[x];
                const vid_t v = hn[u];
                if (u < v && stamp[v] == round && !(v < hn[v])) atomicMin(&claim[v], u);
            }
            __syncthreads();
        }
        for (int32_t x = threadIdx.x; x < len; x += MC_TAIL_BLOCK) {
            const vid_t u = wl[x];
            const vid_t v = hn[u];
            if (!mc_acting(u, v, swap)) continue;
            if (v == u) {
                vcmap[u] = u;
            } else if (stamp[v] == round) {
                if (mc_acting(v, hn[v], swap)) {
                    vcmap[u] = ORD_MAX;
                } else {
                    const vid_t cv = swap ? v : claim[v];
                    vcmap[u] = cv;
                    vcmap[v] = cv;
                }
            } else {
                vcmap[u] = vcmap[v];
            }
        }
        __syncthreads();
        vid_t* next = (wl == wl_a) ? wl_b : wl_a;
        for (int32_t x0 = 0; x0 < len; x0 += MC_TAIL_BLOCK) {
            const int32_t x = x0 + (int32_t)threadIdx.x;
            const vid_t u = x < len ? wl[x] : 0;
            const int32_t at = wave_append(&cnt, x < len && vcmap[u] >= n);
            if (at >= 0) next[at] = u;
        }
        __syncthreads();
        len = cnt;
        wl = next;
        swap ^= 1;
        __syncthreads();
    }
}

__global__ void k_root_flags(const vid_t* vcmap, vid_t n, int32_t* flags) {
    FOR_THREAD(u, n) flags[u] = vcmap[u] == (vid_t)u ? 1 : 0;
}

__global__ void k_apply_ids(const vid_t* vcmap, const int32_t* ids, vid_t n, vid_t* map) {
    FOR_THREAD(u, n) map[u] = ids[vcmap[u]];
}

// Roots (vcmap[u] == u) get dense ids in increasing u; everyone takes its
// root's id (heuristics.hpp:147-165, :650-665).
static int32_t finalize_ids(vid_t n, vid_t* vcmap, vid_t* map, Arena& a, hipStream_t s) {
    ArenaScope sc(a);
    int32_t* ids = a.alloc<int32_t>((int64_t)n + 1);
    int blocks = grid_for(n, 256, 8192);
    k_root_flags<<<blocks, 256, 0, s>>>(vcmap, n, ids);
    JET_LAUNCH_CHECK();
    excl_scan_i32_i32(ids, ids, n, a, s);
    k_apply_ids<<<blocks, 256, 0, s>>>(vcmap, ids, n, map);
    JET_LAUNCH_CHECK();
    return read_scalar(ids + n, s);
}

int32_t map_construct(vid_t n, const vid_t* hn, vid_t* map, Arena& a, hipStream_t s) {
    if (n <= 0) return 0;
    ArenaScope sc(a);
    vid_t* vcmap = a.alloc<vid_t>(n);
    int32_t* stamp = a.alloc<int32_t>(n);
    int32_t* claim = a.alloc<int32_t>(n);
    vid_t* wl_a = a.alloc<vid_t>(n);
    vid_t* wl_b = a.alloc<vid_t>(n);
    int32_t* cnt = a.alloc<int32_t>(1);
    fill_i32(vcmap, ORD_MAX, n, s);
    fill_i32(stamp, -1, n, s);
    const vid_t* wl = nullptr;
    int32_t len = n;
    int swap = 1;
    for (int32_t round = 0; len > 0; round++) {
        if (round > 2 * n + 8) throw DeviceError("map construction made no progress");
        if (wl && len <= MC_TAIL_MAX) {
            // the remaining rounds in one workgroup, with no host read per round
            int32_t* err = a.alloc<int32_t>(1);  // cleared by the kernel
            vid_t* other = (wl == wl_a) ? wl_b : wl_a;
            k_mc_tail<<<1, MC_TAIL_BLOCK, 0, s>>>(wl, len, round, swap, n, hn, stamp, claim, vcmap,
                                                  const_cast<vid_t*>(wl), other, err);
            JET_LAUNCH_CHECK();
            if (read_scalar(err, s)) throw DeviceError("map construction made no progress");
            break;
        }
        int blocks = grid_for(len, 256, 8192);
        k_mc_mark<<<blocks, 256, 0, s>>>(wl, len, round, stamp, claim, cnt);
        if (swap == 0) k_mc_claim<<<blocks, 256, 0, s>>>(wl, len, round, hn, stamp, claim);
        k_mc_resolve<<<blocks, 256, 0, s>>>(wl, len, round, swap, hn, stamp, claim, vcmap);
        vid_t* next = (wl == wl_a) ? wl_b : wl_a;
        k_mc_next<<<blocks, 256, 0, s>>>(wl, len, n, vcmap, next, cnt);
        JET_LAUNCH_CHECK();
        len = read_scalar(cnt, s);
        wl = next;
        swap ^= 1;
    }
    return finalize_ids(n, vcmap, map, a, s);
}

int32_t coarsen_hec(const DGraph& g, bool uniform, uint64_t seed, int32_t level, vid_t* map, Arena& a,
                    hipStream_t s) {
    ArenaScope sc(a);
    vid_t* hn = a.alloc<vid_t>(g.n);
    hec_heavy_neighbours(g, uniform, seed, level, hn, a, s);
    return map_construct(g.n, hn, map, a, s);
}

// ---------------------------------------------------------------- matching
__global__ void k_match_init_uniform(DGraph g, uint64_t seed, uint32_t level, vid_t* hn) {
    FOR_THREAD(i, g.n) {
        eoff_t s0 = g.rowptr[i];
        int64_t deg = g.rowptr[i + 1] - s0;
        if (deg == 0) continue;
        uint64_t r = rng(seed, level, SITE_MATCH_INIT, 0, (uint32_t)i);
        hn[i] = g.col[s0 + (int64_t)(r % (uint64_t)deg)];
    }
}

// pickMatch choice: lexicographic max of (weight, xorshift32(v + r)) over
// eligible neighbours with weight >= 0 (weight ignored when UNIFORM).
__device__ inline void pick_better(int64_t& bw, uint32_t& bh, vid_t& bv, int64_t w, uint32_t h, vid_t v) {
    if (bv == ORD_MAX || w > bw || (w == bw && h > bh)) {
        bw = w;
        bh = h;
        bv = v;
    }
}

template <bool INITIAL>
__device__ inline bool pick_skip(vid_t u, const vid_t* vcmap, const vid_t* hn) {
    if (INITIAL) return false;
    return vcmap[u] != ORD_MAX || hn[u] == ORD_MAX || vcmap[hn[u]] == ORD_MAX;
}

template <int G, bool INITIAL, bool UNIFORM>
__global__ void k_pick(DGraph g, const vid_t* wl, int32_t len, const vid_t* vcmap, vid_t* hn, uint64_t seed,
                       uint32_t level, uint32_t pass) {
    FOR_GROUP(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        eoff_t s0 = g.rowptr[u], e0 = g.rowptr[u + 1];
        if (e0 - s0 > HEAVY_DEG) continue;
        if (pick_skip<INITIAL>(u, vcmap, hn)) continue;
        uint32_t r = (uint32_t)rng(seed, level, SITE_PICK, pass, (uint32_t)u);
        int64_t bw = 0;
        uint32_t bh = 0;
        vid_t bv = ORD_MAX;
        for (eoff_t j = s0 + sub_; j < e0; j += G) {
            vid_t v = g.col[j];
            if (!INITIAL && vcmap[v] != ORD_MAX) continue;
            int64_t w = UNIFORM ? 0 : (int64_t)g.ew[j];
            if (w < 0) continue;
            pick_better(bw, bh, bv, w, xorshift32((uint32_t)v + r), v);
        }
#pragma unroll
        for (int d = G / 2; d >= 1; d >>= 1) {
            int64_t ow = __shfl_xor(bw, d, G);
            uint32_t oh = __shfl_xor(bh, d, G);
            vid_t ov = __shfl_xor(bv, d, G);
            if (ov != ORD_MAX) pick_better(bw, bh, bv, ow, oh, ov);
        }
        if (sub_ == 0) hn[u] = bv;
    }
}

template <bool INITIAL, bool UNIFORM>
__global__ void __launch_bounds__(256) k_pick_heavy(DGraph g, const vid_t* heavy, int32_t nheavy, const vid_t* vcmap,
                                                    vid_t* hn, uint64_t seed, uint32_t level, uint32_t pass) {
    __shared__ int64_t sw[256 / WAVE];
    __shared__ uint32_t sh[256 / WAVE];
    __shared__ vid_t sv[256 / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    for (int32_t t = blockIdx.x; t < nheavy; t += gridDim.x) {
        vid_t u = heavy[t];
        if (pick_skip<INITIAL>(u, vcmap, hn)) continue;
        uint32_t r = (uint32_t)rng(seed, level, SITE_PICK, pass, (uint32_t)u);
        int64_t bw = 0;
        uint32_t bh = 0;
        vid_t bv = ORD_MAX;
        for (eoff_t j = g.rowptr[u] + threadIdx.x; j < g.rowptr[u + 1]; j += blockDim.x) {
            vid_t v = g.col[j];
            if (!INITIAL && vcmap[v] != ORD_MAX) continue;
            int64_t w = UNIFORM ? 0 : (int64_t)g.ew[j];
            if (w < 0) continue;
            pick_better(bw, bh, bv, w, xorshift32((uint32_t)v + r), v);
        }
        for (int d = WAVE / 2; d >= 1; d >>= 1) {
            int64_t ow = __shfl_xor(bw, d, WAVE);
            uint32_t oh = __shfl_xor(bh, d, WAVE);
            vid_t ov = __shfl_xor(bv, d, WAVE);
            if (ov != ORD_MAX) pick_better(bw, bh, bv, ow, oh, ov);
        }
        if (lane == 0) {
            sw[wid] = bw;
            sh[wid] = bh;
            sv[wid] = bv;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < 256 / WAVE; w++)
                if (sv[w] != ORD_MAX) pick_better(bw, bh, bv, sw[w], sh[w], sv[w]);
            hn[u] = bv;
        }
        __syncthreads();
    }
}

template <bool INITIAL, bool UNIFORM>
static void launch_pick(const DGraph& g, const vid_t* wl, int32_t len, const vid_t* vcmap, vid_t* hn, uint64_t seed,
                        int32_t level, uint32_t pass, const vid_t* heavy, int32_t nheavy, hipStream_t s) {
    int G = group_for(g);
    int blocks = grid_for((int64_t)len * G, 256, 16384);
    if (len > 0) {
        switch (G) {
            case 1: k_pick<1, INITIAL, UNIFORM><<<blocks, 256, 0, s>>>(g, wl, len, vcmap, hn, seed, level, pass); break;
            case 8: k_pick<8, INITIAL, UNIFORM><<<blocks, 256, 0, s>>>(g, wl, len, vcmap, hn, seed, level, pass); break;
            case 32: k_pick<32, INITIAL, UNIFORM><<<blocks, 256, 0, s>>>(g, wl, len, vcmap, hn, seed, level, pass); break;
            default: k_pick<64, INITIAL, UNIFORM><<<blocks, 256, 0, s>>>(g, wl, len, vcmap, hn, seed, level, pass); break;
        }
        JET_LAUNCH_CHECK();
    }
    if (nheavy > 0) {
        k_pick_heavy<INITIAL, UNIFORM><<<std::min(nheavy, 1024), 256, 0, s>>>(g, heavy, nheavy, vcmap, hn, seed,
                                                                            (uint32_t)level, pass);
        JET_LAUNCH_CHECK();
    }
}

// Commit round r: reset of round r-1's inactive marks fused with part 1
// (heuristics.hpp:465-499).
__global__ void k_commit_mark(const vid_t* wl, int32_t len, int r, const vid_t* hn, vid_t* vcmap, int32_t* claim) {
    FOR_THREAD(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        vid_t cu = vcmap[u];
        if (cu == ORD_MAX - 1) cu = ORD_MAX;
        claim[u] = ORD_MAX;
        vid_t v = hn[u];
        if (v != ORD_MAX && cu == ORD_MAX) {
            bool cond = r > 0 ? (pod_hash_i32(u + r) < pod_hash_i32(v + r)) : (u < v);
            if (!cond) cu = ORD_MAX - 1;
        }
        vcmap[u] = cu;
    }
}

__global__ void k_commit_claim(const vid_t* wl, int32_t len, const vid_t* hn, const vid_t* vcmap, int32_t* claim) {
    FOR_THREAD(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        vid_t v = hn[u];
        if (v == ORD_MAX || vcmap[u] != ORD_MAX) continue;
        if (vcmap[v] == ORD_MAX - 1) atomicMin(&claim[v], u);
    }
}

__global__ void k_commit_apply(const vid_t* wl, int32_t len, const vid_t* hn, vid_t* vcmap, const int32_t* claim) {
    FOR_THREAD(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        vid_t v = hn[u];
        if (v == ORD_MAX) continue;
        if (claim[v] == u) {
            vid_t cv = u < v ? u : v;
            vcmap[u] = cv;
            vcmap[v] = cv;
        }
    }
}

__global__ void k_commit_reset(const vid_t* wl, int32_t len, vid_t* vcmap) {
    FOR_THREAD(x, len) {
        vid_t u = wl ? wl[x] : (vid_t)x;
        if (vcmap[u] == ORD_MAX - 1) vcmap[u] = ORD_MAX;
    }
}

__global__ void __launch_bounds__(256) k_match_next(const vid_t* wl, int32_t len, const vid_t* vcmap, const vid_t* hn,
                                                    vid_t* next, int32_t* cnt) {
    __shared__ int32_t sh[256 / WAVE + 1];
    FOR_BLOCK_UNIFORM(x, len) {
        vid_t u = 0;
        bool keep = false;
        if (x < len) {
            u = wl ? wl[x] : (vid_t)x;
            keep = vcmap[u] == ORD_MAX && hn[u] != ORD_MAX;
        }
        int32_t p = block_append<256>(cnt, keep, sh);
        if (p >= 0) next[p] = u;
    }
}

__global__ void k_count_unmatched(const vid_t* vcmap, vid_t n, int32_t* cnt) {
    __shared__ int64_t red[256 / WAVE];
    int64_t c = 0;
    FOR_THREAD(i, n) c += vcmap[i] == ORD_MAX ? 1 : 0;
    c = block_reduce_sum<int64_t, 256>(c, red);
    if (threadIdx.x == 0 && c) atomicAdd(cnt, (int32_t)c);
}

// 2-hop candidate flags: stage 0 = leaf (unmatched, degree 1), else unmatched.
__global__ void k_2hop_flags(DGraph g, const vid_t* vcmap, int stage, int32_t* flags) {
    FOR_THREAD(i, g.n) {
        bool f = vcmap[i] == ORD_MAX;
        if (stage == 0) f = f && (g.rowptr[i + 1] - g.rowptr[i] == 1);
        flags[i] = f ? 1 : 0;
    }
}

__global__ void k_2hop_gather(vid_t n, const int32_t* pos, vid_t* cands) {
    FOR_THREAD(i, n) {
        if (pos[i + 1] != pos[i]) cands[pos[i]] = (vid_t)i;
    }
}

// leaf digest: the single neighbour (heuristics.hpp:551-555); relative
// digest: the lowest-degree neighbour, the first heaviest among equals
// (:611-637). Written as uint32 sort keys.
__global__ void k_digest_leaf(DGraph g, const vid_t* cands, int32_t len, uint32_t* keys, uint32_t* vals) {
    FOR_THREAD(x, len) {
        vid_t u = cands[x];
        keys[x] = (uint32_t)g.col[g.rowptr[u]];
        vals[x] = (uint32_t)u;
    }
}

__global__ void k_digest_relative(DGraph g, const vid_t* cands, int32_t len, uint32_t* keys, uint32_t* vals) {
    FOR_THREAD(x, len) {
        vid_t u = cands[x];
        vid_t h = ORD_MAX;
        wgt_t max_wgt = 0;
        vid_t min_deg = ORD_MAX;
        for (eoff_t j = g.rowptr[u]; j < g.rowptr[u + 1]; j++) {
            vid_t v = g.col[j];
            vid_t vd = (vid_t)(g.rowptr[v + 1] - g.rowptr[v]);
            if (min_deg > vd) {
                min_deg = vd;
                max_wgt = g.ew[j];
                h = v;
            } else if (min_deg == vd && max_wgt < g.ew[j]) {
                h = v;
                max_wgt = g.ew[j];
            }
        }
        keys[x] = (uint32_t)h;
        vals[x] = (uint32_t)u;
    }
}

// twin digest: sum of (y*y + y), y = pod_hash(neighbour) widened to 64 bits
// (heuristics.hpp:580-596).
template <int G>
__global__ void k_digest_twin(DGraph g, const vid_t* cands, int32_t len, uint64_t* keys, uint32_t* vals) {
    FOR_GROUP(x, len) {
        vid_t u = cands[x];
        uint64_t h = 0;
        for (eoff_t j = g.rowptr[u] + sub_; j < g.rowptr[u + 1]; j += G) {
            uint64_t y = pod_hash_i32(g.col[j]);
            h += y * y + y;
        }
        h = group_sum<G, uint64_t>(h);
        if (sub_ == 0) {
            keys[x] = h;
            vals[x] = (uint32_t)u;
        }
    }
}

template <class K>
__global__ void k_pair_heads(const K* keys, int32_t len, int32_t* head) {
    FOR_THREAD(t, len) head[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}

__global__ void k_pair_start(const int32_t* head_ex, int32_t len, int32_t* gstart) {
    FOR_THREAD(t, len) {
        if (head_ex[t + 1] != head_ex[t]) gstart[head_ex[t]] = (int32_t)t;
    }
}

__global__ void k_pair_apply(const int32_t* head_ex, const int32_t* gstart, const uint32_t* ids, int32_t len,
                             vid_t* vcmap) {
    FOR_THREAD(t, len) {
        int32_t gid = head_ex[t + 1] - 1;
        int32_t rank = (int32_t)t - gstart[gid];
        if (rank & 1) {
            vid_t a = (vid_t)ids[t - 1], b = (vid_t)ids[t];
            vid_t cv = a < b ? a : b;
            vcmap[a] = cv;
            vcmap[b] = cv;
        }
    }
}

// Groups candidates (given in increasing id) by digest with a stable sort and
// pairs consecutive members of each group.
template <class K>
static void pair_by_digest(K* keys, uint32_t* ids, int32_t len, int key_bits, vid_t* vcmap, Arena& a,
                           hipStream_t s) {
    if (len < 2) return;
    ArenaScope sc(a);
    K* ktmp = a.alloc<K>(len);
    uint32_t* vtmp = a.alloc<uint32_t>(len);
    if (sizeof(K) == 8)
        radix_sort_u64(reinterpret_cast<uint64_t*>(keys), ids, len, key_bits, reinterpret_cast<uint64_t*>(ktmp), vtmp,
                       a, s);
    else
        radix_sort_u32(reinterpret_cast<uint32_t*>(keys), ids, len, key_bits, reinterpret_cast<uint32_t*>(ktmp), vtmp,
                       a, s);
    int32_t* head = a.alloc<int32_t>((int64_t)len + 1);
    int32_t* gstart = a.alloc<int32_t>(len);
    int blocks = grid_for(len, 256, 8192);
    k_pair_heads<K><<<blocks, 256, 0, s>>>(keys, len, head);
    JET_LAUNCH_CHECK();
    excl_scan_i32_i32(head, head, len, a, s);
    k_pair_start<<<blocks, 256, 0, s>>>(head, len, gstart);
    k_pair_apply<<<blocks, 256, 0, s>>>(head, gstart, ids, len, vcmap);
    JET_LAUNCH_CHECK();
}

__global__ void k_singletons(vid_t* vcmap, vid_t n) {
    FOR_THREAD(i, n) {
        if (vcmap[i] == ORD_MAX) vcmap[i] = (vid_t)i;
    }
}

static int32_t count_unmatched(const vid_t* vcmap, vid_t n, Arena& a, hipStream_t s) {
    ArenaScope sc(a);
    int32_t* cnt = a.alloc<int32_t>(1);
    fill_i32(cnt, 0, 1, s);
    k_count_unmatched<<<grid_for(n, 256, 1024), 256, 0, s>>>(vcmap, n, cnt);
    JET_LAUNCH_CHECK();
    return read_scalar(cnt, s);
}

int32_t coarsen_match(const DGraph& g, bool uniform, uint64_t seed, int32_t level, bool two_hop, vid_t* map,
                      Arena& a, hipStream_t s) {
    vid_t n = g.n;
    if (n <= 0) return 0;
    ArenaScope sc(a);
    vid_t* hn = a.alloc<vid_t>(n);
    vid_t* vcmap = a.alloc<vid_t>(n);
    int32_t* claim = a.alloc<int32_t>(n);
    vid_t* wl_a = a.alloc<vid_t>(n);
    vid_t* wl_b = a.alloc<vid_t>(n);
    int32_t* cnt = a.alloc<int32_t>(1);
    vid_t* heavy = a.alloc<vid_t>(n);
    int32_t nheavy = build_heavy_list(g, heavy, a, s);
    fill_i32(hn, ORD_MAX, n, s);
    fill_i32(vcmap, ORD_MAX, n, s);
    if (uniform) {
        k_match_init_uniform<<<grid_for(n, 256, 8192), 256, 0, s>>>(g, seed, (uint32_t)level, hn);
        JET_LAUNCH_CHECK();
    } else {
        launch_pick<true, false>(g, nullptr, n, vcmap, hn, seed, level, 0, heavy, nheavy, s);
    }
    const vid_t* wl = nullptr;
    int32_t len = n;
    uint32_t pass = 0;
    while (len > 0) {
        if (pass > (uint32_t)n + 8) throw DeviceError("matching made no progress");
        int blocks = grid_for(len, 256, 8192);
        for (int r = 0; r < 4; r++) {
            k_commit_mark<<<blocks, 256, 0, s>>>(wl, len, r, hn, vcmap, claim);
            k_commit_claim<<<blocks, 256, 0, s>>>(wl, len, hn, vcmap, claim);
            k_commit_apply<<<blocks, 256, 0, s>>>(wl, len, hn, vcmap, claim);
        }
        k_commit_reset<<<blocks, 256, 0, s>>>(wl, len, vcmap);
        JET_LAUNCH_CHECK();
        pass++;
        if (uniform) launch_pick<false, true>(g, wl, len, vcmap, hn, seed, level, pass, heavy, nheavy, s);
        else launch_pick<false, false>(g, wl, len, vcmap, hn, seed, level, pass, heavy, nheavy, s);
        vid_t* next = (wl == wl_a) ? wl_b : wl_a;
        fill_i32(cnt, 0, 1, s);
        k_match_next<<<blocks, 256, 0, s>>>(wl, len, vcmap, hn, next, cnt);
        JET_LAUNCH_CHECK();
        len = read_scalar(cnt, s);
        wl = next;
    }
    if (two_hop) {
        int32_t* flags = a.alloc<int32_t>((int64_t)n + 1);
        vid_t* cands = a.alloc<vid_t>(n);
        uint64_t* keys64 = a.alloc<uint64_t>(n);
        uint32_t* keys32 = reinterpret_cast<uint32_t*>(keys64);
        uint32_t* ids = a.alloc<uint32_t>(n);
        int vbits = 1;
        while (vbits < 31 && (int64_t(1) << vbits) <= (int64_t)ORD_MAX) vbits++;
        for (int stage = 0; stage < 3; stage++) {
            int32_t unmapped = count_unmatched(vcmap, n, a, s);
            if (!((double)unmapped / (double)n > 0.25)) continue;
            int blocks = grid_for(n, 256, 8192);
            k_2hop_flags<<<blocks, 256, 0, s>>>(g, vcmap, stage, flags);
            JET_LAUNCH_CHECK();
            excl_scan_i32_i32(flags, flags, n, a, s);
            k_2hop_gather<<<blocks, 256, 0, s>>>(n, flags, cands);
            JET_LAUNCH_CHECK();
            int32_t nc = read_scalar(flags + n, s);
            if (nc < 2) continue;
            int cb = grid_for(nc, 256, 8192);
            if (stage == 0) {
                k_digest_leaf<<<cb, 256, 0, s>>>(g, cands, nc, keys32, ids);
                JET_LAUNCH_CHECK();
                pair_by_digest<uint32_t>(keys32, ids, nc, 31, vcmap, a, s);
            } else if (stage == 1) {
                int G = group_for(g);
                int tb = grid_for((int64_t)nc * G, 256, 16384);
                switch (G) {
                    case 1: k_digest_twin<1><<<tb, 256, 0, s>>>(g, cands, nc, keys64, ids); break;
                    case 8: k_digest_twin<8><<<tb, 256, 0, s>>>(g, cands, nc, keys64, ids); break;
                    case 32: k_digest_twin<32><<<tb, 256, 0, s>>>(g, cands, nc, keys64, ids); break;
                    default: k_digest_twin<64><<<tb, 256, 0, s>>>(g, cands, nc, keys64, ids); break;
                }
                JET_LAUNCH_CHECK();
                pair_by_digest<uint64_t>(keys64, ids, nc, 64, vcmap, a, s);
            } else {
                k_digest_relative<<<cb, 256, 0, s>>>(g, cands, nc, keys32, ids);
                JET_LAUNCH_CHECK();
                pair_by_digest<uint32_t>(keys32, ids, nc, 31, vcmap, a, s);
            }
        }
        (void)vbits;
    }
    k_singletons<<<grid_for(n, 256, 8192), 256, 0, s>>>(vcmap, n);
    JET_LAUNCH_CHECK();
    return finalize_ids(n, vcmap, map, a, s);
}

}  // namespace jet
