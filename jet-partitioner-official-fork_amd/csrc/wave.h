// Wave64 / workgroup reduction and scan helpers (device only).
#pragma once
#include <hip/hip_runtime.h>

#include "jet_types.h"

namespace jet {

// Grid-stride loop, one item per thread.
#define FOR_THREAD(IDX_, LEN_)                                                                \
    for (int64_t IDX_ = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; IDX_ < (LEN_);         \
         IDX_ += (int64_t)gridDim.x * blockDim.x)

// One contiguous range of items per workgroup (ranges rounded to blockDim),
// one item per thread and pass: consecutive items share a workgroup, so a
// workgroup-local aggregation sees fewer distinct keys than a grid stride.
#define FOR_BLOCK_RANGE(IDX_, LEN_)                                                                  \
    const int64_t per_block_ =                                                                       \
        (((int64_t)(LEN_) + gridDim.x - 1) / gridDim.x + blockDim.x - 1) / blockDim.x * blockDim.x;   \
    const int64_t range_end_ = min((int64_t)(LEN_), (int64_t)(blockIdx.x + 1) * per_block_);         \
    for (int64_t IDX_ = (int64_t)blockIdx.x * per_block_ + threadIdx.x; IDX_ < range_end_; IDX_ += blockDim.x)

// Grid-stride loop, one item per group of G lanes (G a power of two <= 64);
// defines sub_ = lane index inside the group.
#define FOR_GROUP(IDX_, LEN_)                                                                 \
    const int sub_ = threadIdx.x & (G - 1);                                                   \
    const int64_t gstride_ = ((int64_t)gridDim.x * blockDim.x) / G;                           \
    for (int64_t IDX_ = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; IDX_ < (LEN_); IDX_ += gstride_)

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Appends the lanes with pred set to a list guarded by *counter: one atomic
// per wave; returns each such lane's slot (or -1). Call from every active lane.
__device__ inline int32_t wave_append(int32_t* counter, bool pred) {
    const int lane = threadIdx.x & (WAVE - 1);
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int leader = __ffsll((unsigned long long)mask) - 1;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (int32_t)__popcll(mask));
    base = __shfl(base, leader, WAVE);
    return pred ? base + (int32_t)__popcll(mask & ((1ull << lane) - 1ull)) : -1;
}

// Workgroup-aggregated append: one atomic per workgroup and call instead of
// one per wave (thousands of waves adding to one counter serialise on it).
// Every thread of the workgroup must make the same sequence of calls; sh
// holds BLOCK / WAVE + 1 ints. Returns the slot of each pred lane, else -1.
template <int BLOCK>
__device__ inline int32_t block_append(int32_t* counter, bool pred, int32_t* sh) {
    constexpr int WAVES = BLOCK / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const uint64_t mask = __ballot(pred);
    if (lane == 0) sh[w] = (int32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t t = 0;
        for (int q = 0; q < WAVES; q++) {
            const int32_t c = sh[q];
            sh[q] = t;
            t += c;
        }
        sh[WAVES] = t ? atomicAdd(counter, t) : 0;
    }
    __syncthreads();
    const int32_t r = pred ? sh[WAVES] + sh[w] + (int32_t)__popcll(mask & ((1ull << lane) - 1ull)) : -1;
    __syncthreads();
    return r;
}

// Grid-stride loop whose trip count is the same for every thread of a
// workgroup (IDX_ may run past LEN_: check it), so the body may call
// block_append.
#define FOR_BLOCK_UNIFORM(IDX_, LEN_)                                                              \
    for (int64_t base_ = (int64_t)blockIdx.x * blockDim.x, IDX_ = base_ + threadIdx.x; base_ < (LEN_); \
         base_ += (int64_t)gridDim.x * blockDim.x, IDX_ = base_ + threadIdx.x)

// Unordered two-list append with wave-private LDS staging and one global
// atomic per workgroup and list. Each wave stages up to CHW items (front list
// grows up, back list grows down inside its region); flush() publishes them:
// front items to list[base..], back items to list[nmax-1-base..].
template <int WAVES, int CHW>
struct StagedLists {
    vid_t* buf;  // LDS, WAVES * CHW
    int32_t nf = 0, nb = 0;
    __device__ explicit StagedLists(vid_t* lds) : buf(lds + (threadIdx.x / WAVE) * CHW) {}
    __device__ inline void push(bool pf, bool pb, vid_t v) {
        const int lane = threadIdx.x & (WAVE - 1);
        const uint64_t lt = (1ull << lane) - 1ull;
        uint64_t mf = __ballot(pf), mb = __ballot(pb);
        if (pf) buf[nf + (int32_t)__popcll(mf & lt)] = v;
        if (pb) buf[CHW - 1 - (nb + (int32_t)__popcll(mb & lt))] = v;
        nf += (int32_t)__popcll(mf);
        nb += (int32_t)__popcll(mb);
    }
    // All threads of the workgroup must call flush().
    __device__ inline void flush(vid_t* list, int32_t* cnt_front, int32_t* cnt_back, vid_t nmax, int32_t* sh) {
        const int lane = threadIdx.x & (WAVE - 1);
        const int w = threadIdx.x / WAVE;
        if (lane == 0) {
            sh[2 * w] = nf;
            sh[2 * w + 1] = nb;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int32_t tf = 0, tb = 0;
            for (int q = 0; q < WAVES; q++) {
                int32_t a = sh[2 * q], b = sh[2 * q + 1];
                sh[2 * q] = tf;
                sh[2 * q + 1] = tb;
                tf += a;
                tb += b;
            }
            sh[2 * WAVES] = tf ? atomicAdd(cnt_front, tf) : 0;
            sh[2 * WAVES + 1] = tb ? atomicAdd(cnt_back, tb) : 0;
        }
        __syncthreads();
        int32_t bf = sh[2 * WAVES] + sh[2 * w], bb = sh[2 * WAVES + 1] + sh[2 * w + 1];
        for (int32_t t = lane; t < nf; t += WAVE) list[bf + t] = buf[t];
        for (int32_t t = lane; t < nb; t += WAVE) list[nmax - 1 - (bb + t)] = buf[CHW - 1 - t];
    }
    // flush() that also writes stamp[v] = gen for every listed v
    __device__ inline void flush_stamped(vid_t* list, int32_t* cnt_front, int32_t* cnt_back, vid_t nmax, int32_t* sh,
                                         int32_t* stamp, int32_t gen) {
        flush(list, cnt_front, cnt_back, nmax, sh);
        const int lane = threadIdx.x & (WAVE - 1);
        for (int32_t t = lane; t < nf; t += WAVE) stamp[buf[t]] = gen;
        for (int32_t t = lane; t < nb; t += WAVE) stamp[buf[CHW - 1 - t]] = gen;
    }
};

template <class T>
__device__ inline T wave_incl_scan(T x) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    return x;
}

template <class T>
__device__ inline T wave_reduce_sum(T x) {
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

template <class T>
__device__ inline T wave_reduce_max(T x) {
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
        T y = __shfl_xor(x, d, WAVE);
        x = y > x ? y : x;
    }
    return x;
}

// Sum over the workgroup; every thread gets the result. lds: BLOCK/64 slots.
template <class T, int BLOCK>
__device__ inline T block_reduce_sum(T x, T* lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    x = wave_reduce_sum(x);
    if (lane == 0) lds[wid] = x;
    __syncthreads();
    T t = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / WAVE; w++) t += lds[w];
    __syncthreads();
    return t;
}

template <class T, int BLOCK>
__device__ inline T block_reduce_max(T x, T* lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    x = wave_reduce_max(x);
    if (lane == 0) lds[wid] = x;
    __syncthreads();
    T t = lds[0];
#pragma unroll
    for (int w = 1; w < BLOCK / WAVE; w++) t = lds[w] > t ? lds[w] : t;
    __syncthreads();
    return t;
}

// Exclusive scan over the workgroup. lds: BLOCK/64 + 1 slots.
template <class T, int BLOCK>
__device__ inline T block_excl_scan(T x, T& total, T* lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    T inc = wave_incl_scan(x);
    if (lane == WAVE - 1) lds[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int w = 0; w < BLOCK / WAVE; w++) {
            T t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[BLOCK / WAVE] = run;
    }
    __syncthreads();
    T res = inc - x + lds[wid];
    total = lds[BLOCK / WAVE];
    __syncthreads();
    return res;
}

// Largest r in [lo, hi] with rowptr[r] <= e: the row holding edge e when
// rowptr[lo] <= e < rowptr[hi + 1].
__device__ inline vid_t row_of(const eoff_t* rowptr, vid_t lo, vid_t hi, eoff_t e) {
    while (lo < hi) {
        vid_t mid = lo + (hi - lo + 1) / 2;
        if (rowptr[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Edge-balanced CSR walk, independent of the degree distribution: workgroups
// take tiles of EB_TILE consecutive edges (grid-stride) and find each edge's
// row by binary search over the rows the tile spans, held in LDS when at most
// EB_SEG of them (a tile over many empty rows searches global memory). Calls
// f(row, edge) for every edge. All threads of the workgroup must call it.
constexpr int EB_TILE = 2048;
constexpr int EB_SEG = 1024;
struct EdgeWalkLds {
    eoff_t seg[EB_SEG];
    vid_t r[2];
};
template <int BLOCK, class F>
__device__ inline void block_edges(const eoff_t* rowptr, vid_t n, eoff_t m, EdgeWalkLds& L, F&& f) {
    for (int64_t t0 = blockIdx.x * (int64_t)EB_TILE; t0 < m; t0 += (int64_t)gridDim.x * EB_TILE) {
        const int64_t t1 = t0 + EB_TILE < m ? t0 + EB_TILE : m;
        if (threadIdx.x == 0) {
            L.r[0] = row_of(rowptr, 0, n - 1, t0);
            L.r[1] = row_of(rowptr, L.r[0], n - 1, t1 - 1);
        }
        __syncthreads();
        const vid_t r0 = L.r[0], r1 = L.r[1];
        const int32_t span = r1 - r0 + 1;
        const bool in_lds = span <= EB_SEG;
        if (in_lds)
            for (int q = threadIdx.x; q < span; q += BLOCK) L.seg[q] = rowptr[r0 + q];
        __syncthreads();
        for (int64_t e = t0 + threadIdx.x; e < t1; e += BLOCK) {
            vid_t r;
            if (in_lds) {
                int32_t lo = 0, hi = span - 1;
                while (lo < hi) {
                    int32_t mid = (lo + hi + 1) >> 1;
                    if (L.seg[mid] <= e) lo = mid;
                    else hi = mid - 1;
                }
                r = r0 + lo;
            } else {
                r = row_of(rowptr, r0, r1, e);
            }
            f(r, e);
        }
        __syncthreads();
    }
}

}  // namespace jet
