// Device-wide primitives for the Jet hot path on gfx950.
#include "prims.h"

#include <algorithm>
#include <stdexcept>
#include "wave.h"

namespace jet {

// ---------------------------------------------------------------- fills
__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill2_i32(int32_t* p1, int32_t v1, int64_t n1, int32_t* p2, int32_t v2, int64_t n2) {
    const int64_t n = n1 > n2 ? n1 : n2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (i < n1) p1[i] = v1;
        if (i < n2) p2[i] = v2;
    }
}
__global__ void k_fill_i64(int64_t* p, int64_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_iota_i32(int32_t* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}
__global__ void k_copy_i32(int32_t* d, const int32_t* s, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

__global__ void k_copy_i64(int64_t* d, const int64_t* s, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}
__global__ void k_peek(const unsigned char* src, int bytes, int64_t* dst) {
    int64_t v = 0;
    for (int b = 0; b < bytes; b++) v |= (int64_t)src[b] << (8 * b);
    *dst = v;
    __threadfence_system();
}

Mailbox::~Mailbox() {
    if (host) (void)hipHostFree(host);
}

Mailbox& thread_mailbox() {
    static thread_local Mailbox mb;
    if (!mb.host) {
        JET_HIP(hipHostMalloc(reinterpret_cast<void**>(&mb.host), MAILBOX_SLOTS * sizeof(int64_t),
                              hipHostMallocMapped | hipHostMallocCoherent));
        JET_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&mb.dev), mb.host, 0));
    }
    return mb;
}

void peek_to_mailbox(const void* src, int bytes, int slot, hipStream_t s) {
    Mailbox& mb = thread_mailbox();
    k_peek<<<1, 1, 0, s>>>(static_cast<const unsigned char*>(src), bytes, mb.dev + slot);
    JET_LAUNCH_CHECK();
}

__global__ void k_peek_words(const int32_t* src, int count, int64_t* dst) {
    if (threadIdx.x < count) dst[threadIdx.x] = src[threadIdx.x];
    __threadfence_system();
}

void read_i32_words(const int32_t* src, int count, int32_t* out, hipStream_t s) {
    if (count <= 0) return;
    if (count > 8) throw std::invalid_argument("read_i32_words: at most 8 words (mailbox slots 0-7)");
    Mailbox& mb = thread_mailbox();
    k_peek_words<<<1, 64, 0, s>>>(src, count, mb.dev);
    JET_LAUNCH_CHECK();
    JET_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < count; i++) out[i] = (int32_t)((volatile int64_t*)mb.host)[i];
}

void copy_i64(int64_t* d, const int64_t* src, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_copy_i64<<<grid_for(n, 256, 8192), 256, 0, s>>>(d, src, n);
    JET_LAUNCH_CHECK();
}

void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_fill_i32<<<grid_for(n, 256, 8192), 256, 0, s>>>(p, v, n);
    JET_LAUNCH_CHECK();
}
void fill2_i32(int32_t* p1, int32_t v1, int64_t n1, int32_t* p2, int32_t v2, int64_t n2, hipStream_t s) {
    const int64_t n = std::max<int64_t>(n1, n2);
    if (n <= 0) return;
    k_fill2_i32<<<grid_for(n, 256, 8192), 256, 0, s>>>(p1, v1, n1, p2, v2, n2);
    JET_LAUNCH_CHECK();
}
void fill_i64(int64_t* p, int64_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_fill_i64<<<grid_for(n, 256, 8192), 256, 0, s>>>(p, v, n);
    JET_LAUNCH_CHECK();
}
void iota_i32(int32_t* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_iota_i32<<<grid_for(n, 256, 8192), 256, 0, s>>>(p, n);
    JET_LAUNCH_CHECK();
}
void copy_i32(int32_t* d, const int32_t* src, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_copy_i32<<<grid_for(n, 256, 8192), 256, 0, s>>>(d, src, n);
    JET_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- scans
constexpr int SB = 256;
constexpr int SI = 16;
constexpr int64_t STILE = (int64_t)SB * SI;

template <class Tin>
__global__ void __launch_bounds__(SB) k_scan_reduce(const Tin* in, int64_t n, int64_t* bsum) {
    __shared__ int64_t red[SB / WAVE];
    int64_t base = blockIdx.x * STILE;
    int64_t s = 0;
    for (int i = 0; i < SI; i++) {
        int64_t idx = base + (int64_t)i * SB + threadIdx.x;
        if (idx < n) s += (int64_t)in[idx];
    }
    s = block_reduce_sum<int64_t, SB>(s, red);
    if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

// Tile of STILE items: coalesced load into LDS (padded against bank
// conflicts), each thread scans SI consecutive items, one workgroup scan of
// the thread totals, coalesced store.
__device__ inline int pad_idx(int li) { return li + (li >> 5); }

template <class Tin, class Tout>
__global__ void __launch_bounds__(SB) k_scan_down(const Tin* in, Tout* out, int64_t n, const int64_t* boff) {
    __shared__ int64_t tile[STILE + STILE / 32];
    __shared__ int64_t ws[SB / WAVE + 1];
    const int64_t base = blockIdx.x * STILE;
    for (int i = 0; i < SI; i++) {
        int li = i * SB + threadIdx.x;
        int64_t idx = base + li;
        tile[pad_idx(li)] = idx < n ? (int64_t)in[idx] : 0;
    }
    __syncthreads();
    int64_t sum = 0;
#pragma unroll
    for (int j = 0; j < SI; j++) sum += tile[pad_idx(threadIdx.x * SI + j)];
    int64_t total;
    int64_t ex = block_excl_scan<int64_t, SB>(sum, total, ws);
    const int64_t b0 = boff ? boff[blockIdx.x] : 0;
    int64_t carry = b0 + ex;
#pragma unroll
    for (int j = 0; j < SI; j++) {
        int li = pad_idx(threadIdx.x * SI + j);
        int64_t v = tile[li];
        tile[li] = carry;
        carry += v;
    }
    __syncthreads();
    for (int i = 0; i < SI; i++) {
        int li = i * SB + threadIdx.x;
        int64_t idx = base + li;
        if (idx < n) out[idx] = (Tout)tile[pad_idx(li)];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = (Tout)(b0 + total);
}

// Arrays of up to SC_MAX items (the rebalancers' bucket sums and the
// table offsets of mid-sized levels): one workgroup walks the array in tiles
// of SC_TILE items. A tile is loaded coalesced into LDS (padded against bank
// conflicts, 132 KB of gfx950's 160 KB), each thread scans SC_PER consecutive
// items, one workgroup scan of the thread totals, then the tile is stored
// coalesced. One launch instead of three; in place is safe, as a tile is
// read whole before it is written and later tiles lie beyond it.
constexpr int SC_BLOCK = 1024;
constexpr int SC_PER = 16;
constexpr int SC_TILE = SC_BLOCK * SC_PER;
constexpr int64_t SC_MAX = 2 * (int64_t)SC_TILE;

template <class Tin, class Tout>
__global__ void __launch_bounds__(SC_BLOCK) k_scan_chunked(const Tin* in, Tout* out, int64_t n) {
    __shared__ int64_t tile[SC_TILE + SC_TILE / 32];
    __shared__ int64_t ws[SC_BLOCK / WAVE + 1];
    int64_t carry_in = 0;
    for (int64_t base = 0; base < n; base += SC_TILE) {
#pragma unroll
        for (int i = 0; i < SC_PER; i++) {
            const int li = i * SC_BLOCK + threadIdx.x;
            const int64_t idx = base + li;
            tile[pad_idx(li)] = idx < n ? (int64_t)in[idx] : 0;
        }
        __syncthreads();
        int64_t sum = 0;
#pragma unroll
        for (int j = 0; j < SC_PER; j++) sum += tile[pad_idx(threadIdx.x * SC_PER + j)];
        int64_t total;
        int64_t carry = carry_in + block_excl_scan<int64_t, SC_BLOCK>(sum, total, ws);
#pragma unroll
        for (int j = 0; j < SC_PER; j++) {
            const int li = pad_idx(threadIdx.x * SC_PER + j);
            const int64_t v = tile[li];
            tile[li] = carry;
            carry += v;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SC_PER; i++) {
            const int li = i * SC_BLOCK + threadIdx.x;
            const int64_t idx = base + li;
            if (idx < n) out[idx] = (Tout)tile[pad_idx(li)];
        }
        carry_in += total;
        __syncthreads();  // the tile and ws are reused by the next tile
    }
    if (threadIdx.x == 0) out[n] = (Tout)carry_in;
}

// Out-of-place arrays of up to SC_MAX items in one launch and without waits
// between workgroups: workgroup b scans tile b (SP_TILE items, coalesced
// through LDS) and sums every item before the tile itself (at most SC_MAX
// items, read from L2 by all its threads at once). Not for in == out: a
// workgroup would read a tile that another has already overwritten.
constexpr int SP_BLOCK = 1024;
constexpr int SP_PER = 4;
constexpr int SP_TILE = SP_BLOCK * SP_PER;

template <class Tin, class Tout>
__global__ void __launch_bounds__(SP_BLOCK) k_scan_prefix_tiles(const Tin* __restrict__ in, Tout* __restrict__ out,
                                                                int64_t n) {
    __shared__ int64_t tile[SP_TILE + SP_TILE / 32];
    __shared__ int64_t ws[SP_BLOCK / WAVE + 1];
    __shared__ int64_t red[SP_BLOCK / WAVE];
    const int64_t base = (int64_t)blockIdx.x * SP_TILE;
    int64_t v[SP_PER];
#pragma unroll
    for (int i = 0; i < SP_PER; i++) {
        const int64_t idx = base + i * SP_BLOCK + threadIdx.x;
        v[i] = idx < n ? (int64_t)in[idx] : 0;
    }
    int64_t before = 0;
    for (int64_t j0 = 0; j0 < base; j0 += 8 * SP_BLOCK) {
        int64_t t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int64_t idx = j0 + u * SP_BLOCK + threadIdx.x;
            t[u] = idx < base ? (int64_t)in[idx] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) before += t[u];
    }
    before = block_reduce_sum<int64_t, SP_BLOCK>(before, red);
#pragma unroll
    for (int i = 0; i < SP_PER; i++) tile[pad_idx(i * SP_BLOCK + threadIdx.x)] = v[i];
    __syncthreads();
    int64_t sum = 0;
#pragma unroll
    for (int j = 0; j < SP_PER; j++) sum += tile[pad_idx(threadIdx.x * SP_PER + j)];
    int64_t total;
    int64_t carry = before + block_excl_scan<int64_t, SP_BLOCK>(sum, total, ws);
#pragma unroll
    for (int j = 0; j < SP_PER; j++) {
        const int li = pad_idx(threadIdx.x * SP_PER + j);
        const int64_t x = tile[li];
        tile[li] = carry;
        carry += x;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SP_PER; i++) {
        const int li = i * SP_BLOCK + threadIdx.x;
        const int64_t idx = base + li;
        if (idx < n) out[idx] = (Tout)tile[pad_idx(li)];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = (Tout)(before + total);
}

template <class Tin, class Tout>
static void excl_scan_impl(const Tin* in, Tout* out, int64_t n, Arena& a, hipStream_t s) {
    if (n <= 0) {
        if (sizeof(Tout) == 8) fill_i64(reinterpret_cast<int64_t*>(out), 0, 1, s);
        else fill_i32(reinterpret_cast<int32_t*>(out), 0, 1, s);
        return;
    }
    int64_t nb = ceil_div(n, STILE);
    if (nb == 1) {
        k_scan_down<Tin, Tout><<<1, SB, 0, s>>>(in, out, n, nullptr);
        JET_LAUNCH_CHECK();
        return;
    }
    if (n <= SC_MAX && (const void*)in != (const void*)out) {
        k_scan_prefix_tiles<Tin, Tout><<<(unsigned)ceil_div(n, (int64_t)SP_TILE), SP_BLOCK, 0, s>>>(in, out, n);
        JET_LAUNCH_CHECK();
        return;
    }
    if (n <= SC_MAX) {
        k_scan_chunked<Tin, Tout><<<1, SC_BLOCK, 0, s>>>(in, out, n);
        JET_LAUNCH_CHECK();
        return;
    }
    ArenaScope scope(a);
    int64_t* bsum = a.alloc<int64_t>(nb);
    int64_t* boff = a.alloc<int64_t>(nb + 1);
    k_scan_reduce<Tin><<<(unsigned)nb, SB, 0, s>>>(in, n, bsum);
    JET_LAUNCH_CHECK();
    excl_scan_impl<int64_t, int64_t>(bsum, boff, nb, a, s);
    k_scan_down<Tin, Tout><<<(unsigned)nb, SB, 0, s>>>(in, out, n, boff);
    JET_LAUNCH_CHECK();
}

void excl_scan_i32_i32(const int32_t* in, int32_t* out, int64_t n, Arena& a, hipStream_t s) {
    excl_scan_impl<int32_t, int32_t>(in, out, n, a, s);
}
void excl_scan_i32_i64(const int32_t* in, int64_t* out, int64_t n, Arena& a, hipStream_t s) {
    excl_scan_impl<int32_t, int64_t>(in, out, n, a, s);
}
void excl_scan_i64_i64(const int64_t* in, int64_t* out, int64_t n, Arena& a, hipStream_t s) {
    excl_scan_impl<int64_t, int64_t>(in, out, n, a, s);
}

// ---------------------------------------------------------------- radix sort
constexpr int RB = 256;
constexpr int RI = 16;
constexpr int64_t RTILE = (int64_t)RB * RI;
constexpr int RWAVES = RB / WAVE;

template <class K>
__global__ void __launch_bounds__(RB) k_rs_hist(const K* keys, int64_t n, int shift, int32_t* hist, int64_t nblocks) {
    __shared__ int32_t h[256];
    for (int t = threadIdx.x; t < 256; t += RB) h[t] = 0;
    __syncthreads();
    int64_t base = blockIdx.x * RTILE;
    for (int i = 0; i < RI; i++) {
        int64_t idx = base + (int64_t)i * RB + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 256; t += RB) hist[(int64_t)t * nblocks + blockIdx.x] = h[t];
}

// Stable scatter: inside a tile, items are ranked chunk by chunk (RB items),
// wave by wave and lane by lane, so equal digits keep their input order.
template <class K>
__global__ void __launch_bounds__(RB) k_rs_scatter(const K* keys, const uint32_t* vals, int64_t n, int shift,
                                                   const int64_t* offs, int64_t nblocks, K* kout, uint32_t* vout) {
    __shared__ int64_t base_off[256];
    __shared__ int32_t wcnt[RWAVES][256];
    __shared__ int32_t ctot[256];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    for (int t = threadIdx.x; t < 256; t += RB) base_off[t] = offs[(int64_t)t * nblocks + blockIdx.x];
    int64_t base = blockIdx.x * RTILE;
    for (int i = 0; i < RI; i++) {
        for (int t = threadIdx.x; t < RWAVES * 256; t += RB) (&wcnt[0][0])[t] = 0;
        __syncthreads();
        int64_t idx = base + (int64_t)i * RB + threadIdx.x;
        bool valid = idx < n;
        K key = valid ? keys[idx] : K(0);
        uint32_t val = valid ? vals[idx] : 0u;
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t same = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(valid && bit);
            same &= bit ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        uint32_t cnt = (uint32_t)__popcll(same);
        if (valid && rank == cnt - 1) wcnt[wid][d] = (int32_t)cnt;
        __syncthreads();
        for (int t = threadIdx.x; t < 256; t += RB) {
            int32_t run = 0;
            for (int w = 0; w < RWAVES; w++) {
                int32_t c = wcnt[w][t];
                wcnt[w][t] = run;
                run += c;
            }
            ctot[t] = run;
        }
        __syncthreads();
        if (valid) {
            int64_t pos = base_off[d] + wcnt[wid][d] + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < 256; t += RB) base_off[t] += ctot[t];
        __syncthreads();
    }
}

template <class K>
static void radix_sort_impl(K* keys, uint32_t* vals, int64_t n, int key_bits, K* ktmp, uint32_t* vtmp, Arena& a,
                            hipStream_t s) {
    if (n <= 1 || key_bits <= 0) return;
    ArenaScope scope(a);
    int64_t nb = ceil_div(n, RTILE);
    int32_t* hist = a.alloc<int32_t>(256 * nb);
    int64_t* offs = a.alloc<int64_t>(256 * nb + 1);
    K* src_k = keys;
    uint32_t* src_v = vals;
    K* dst_k = ktmp;
    uint32_t* dst_v = vtmp;
    int passes = (key_bits + 7) / 8;
    for (int p = 0; p < passes; p++) {
        int shift = 8 * p;
        k_rs_hist<K><<<(unsigned)nb, RB, 0, s>>>(src_k, n, shift, hist, nb);
        JET_LAUNCH_CHECK();
        excl_scan_i32_i64(hist, offs, 256 * nb, a, s);
        k_rs_scatter<K><<<(unsigned)nb, RB, 0, s>>>(src_k, src_v, n, shift, offs, nb, dst_k, dst_v);
        JET_LAUNCH_CHECK();
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    if (src_k != keys) {
        JET_HIP(hipMemcpyAsync(keys, src_k, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
        JET_HIP(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    }
}

void radix_sort_u32(uint32_t* keys, uint32_t* vals, int64_t n, int key_bits, uint32_t* ktmp, uint32_t* vtmp,
                    Arena& a, hipStream_t s) {
    radix_sort_impl<uint32_t>(keys, vals, n, key_bits, ktmp, vtmp, a, s);
}
void radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int key_bits, uint64_t* ktmp, uint32_t* vtmp,
                    Arena& a, hipStream_t s) {
    radix_sort_impl<uint64_t>(keys, vals, n, key_bits, ktmp, vtmp, a, s);
}

}  // namespace jet
