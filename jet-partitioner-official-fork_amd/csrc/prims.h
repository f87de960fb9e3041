// Device-wide primitives: fills, exclusive scans, stable LSD radix sort.
#pragma once
#include "arena.h"
#include "jet_types.h"

namespace jet {

void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
void fill_i64(int64_t* p, int64_t v, int64_t n, hipStream_t s);
void iota_i32(int32_t* p, int64_t n, hipStream_t s);
void copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s);

// out[i] = in[0] + ... + in[i-1] for i in [0, n]; out has n + 1 entries
// (out[n] is the total). `in` and `out` may alias when the types match.
void excl_scan_i32_i32(const int32_t* in, int32_t* out, int64_t n, Arena& a, hipStream_t s);
void excl_scan_i32_i64(const int32_t* in, int64_t* out, int64_t n, Arena& a, hipStream_t s);
void excl_scan_i64_i64(const int64_t* in, int64_t* out, int64_t n, Arena& a, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on the low `key_bits` bits.
// The sorted pairs are returned in (keys, vals); ktmp/vtmp are n-sized.
void radix_sort_u32(uint32_t* keys, uint32_t* vals, int64_t n, int key_bits, uint32_t* ktmp, uint32_t* vtmp,
                    Arena& a, hipStream_t s);
void radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int key_bits, uint64_t* ktmp, uint32_t* vtmp,
                    Arena& a, hipStream_t s);

inline int grid_for(int64_t n, int block, int max_blocks = 65536) {
    int64_t b = ceil_div(n, block);
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (int)b;
}

// Reads one value from device memory into host memory, synchronously.
template <class T>
T read_scalar(const T* dev, hipStream_t s) {
    T h;
    JET_HIP(hipMemcpyAsync(&h, dev, sizeof(T), hipMemcpyDeviceToHost, s));
    JET_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace jet
