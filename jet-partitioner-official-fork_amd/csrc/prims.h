// Device-wide primitives: fills, exclusive scans, stable LSD radix sort.
#pragma once
#include <cstring>

#include "arena.h"
#include "jet_types.h"

namespace jet {

void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
// Two fills in one launch: p1[0, n1) = v1 and p2[0, n2) = v2.
void fill2_i32(int32_t* p1, int32_t v1, int64_t n1, int32_t* p2, int32_t v2, int64_t n2, hipStream_t s);
void fill_i64(int64_t* p, int64_t v, int64_t n, hipStream_t s);
void iota_i32(int32_t* p, int64_t n, hipStream_t s);
void copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s);
void copy_i64(int64_t* dst, const int64_t* src, int64_t n, hipStream_t s);

// Pinned, device-mapped host words owned by the calling thread (allocated on
// first use, MAILBOX_SLOTS of them). A kernel stores a result there with a
// system-scope fence and the host reads it after synchronising the stream:
// no copy command and no staging buffer on the way.
constexpr int MAILBOX_SLOTS = 64;
struct Mailbox {
    int64_t* host = nullptr;
    int64_t* dev = nullptr;
    // last sequence word a kernel was asked to write here: unique per
    // mailbox, so a stale word never matches a new wait
    int64_t seq = 0;
    ~Mailbox();
};
Mailbox& thread_mailbox();
// Copies `bytes` (<= 8) from device memory into mailbox slot `slot` (stream-ordered).
void peek_to_mailbox(const void* src, int bytes, int slot, hipStream_t s);
// Reads `count` (<= 8) int32 words from device memory, synchronously. Slots
// 0-7 serve these reads and read_scalar; slots 8-15 carry the refiner's
// per-iteration results.
void read_i32_words(const int32_t* src, int count, int32_t* out, hipStream_t s);

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

// Reads one value from device memory into host memory, synchronously,
// through this thread's mailbox.
template <class T>
T read_scalar(const T* dev, hipStream_t s) {
    static_assert(sizeof(T) <= 8, "read_scalar reads at most 8 bytes");
    peek_to_mailbox(dev, (int)sizeof(T), 0, s);
    JET_HIP(hipStreamSynchronize(s));
    T h;
    std::memcpy(&h, (const void*)thread_mailbox().host, sizeof(T));
    return h;
}

}  // namespace jet
