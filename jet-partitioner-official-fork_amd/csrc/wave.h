// Wave64 / workgroup reduction and scan helpers (device only).
#pragma once
#include <hip/hip_runtime.h>

#include "jet_types.h"

namespace jet {

// Grid-stride loop, one item per thread.
#define FOR_THREAD(IDX_, LEN_)                                                                \
    for (int64_t IDX_ = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; IDX_ < (LEN_);         \
         IDX_ += (int64_t)gridDim.x * blockDim.x)

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

}  // namespace jet
