// Input staging on the device: validation, offset widening, weight narrowing,
// table capacity, and a batch evaluator of gain_bucket.
#include "misc.h"
#include "prims.h"
#include "wave.h"

namespace jet {

// Offsets first: monotone and inside [0, m]; only then are columns read.
__global__ void k_validate_offsets(vid_t n, const eoff_t* rowptr, eoff_t m, int32_t* err) {
    int32_t e = 0;
    FOR_THREAD(i, n) {
        eoff_t a = rowptr[i], b = rowptr[i + 1];
        if (b < a || a < 0 || b > m) e |= 1;
    }
    if (e) atomicOr(err, e);
}

// err bits: 2 column out of range, 4 non-positive edge weight, 8 non-positive
// vertex weight, 16 weight total overflows int32 (set on the host). Flat
// passes over vertices and over edges, so hubs do not serialise a thread.
__global__ void k_validate(vid_t n, eoff_t m, const vid_t* col, const wgt_t* ew, const wgt_t* vw, int32_t* err,
                           unsigned long long* sums) {
    __shared__ int64_t red[256 / WAVE];
    int64_t se = 0, sv = 0;
    int32_t e = 0;
    FOR_THREAD(i, n) {
        if (vw) {
            if (vw[i] <= 0) e |= 8;
            sv += vw[i];
        } else {
            sv += 1;
        }
    }
    FOR_THREAD(j, m) {
        vid_t c = col[j];
        if (c < 0 || c >= n) e |= 2;
        if (ew) {
            if (ew[j] <= 0) e |= 4;
            se += ew[j];
        } else {
            se += 1;
        }
    }
    if (e) atomicOr(err, e);
    se = block_reduce_sum<int64_t, 256>(se, red);
    sv = block_reduce_sum<int64_t, 256>(sv, red);
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], (unsigned long long)se);
        atomicAdd(&sums[1], (unsigned long long)sv);
    }
}

void launch_validate_offsets(vid_t n, const eoff_t* rowptr, eoff_t m, int32_t* err, hipStream_t s) {
    k_validate_offsets<<<grid_for(n, 256, 4096), 256, 0, s>>>(n, rowptr, m, err);
    JET_LAUNCH_CHECK();
}

void launch_validate(vid_t n, eoff_t m, const vid_t* col, const wgt_t* ew, const wgt_t* vw, int32_t* err,
                     unsigned long long* sums, hipStream_t s) {
    k_validate<<<grid_for(std::max<int64_t>(n, m), 256, 4096), 256, 0, s>>>(n, m, col, ew, vw, err, sums);
    JET_LAUNCH_CHECK();
}

int32_t validate_graph(vid_t n, const eoff_t* rowptr, const vid_t* col, const wgt_t* ew, const wgt_t* vw, Arena& a,
                       hipStream_t s) {
    ArenaScope sc(a);
    int32_t* err = a.alloc<int32_t>(1);
    unsigned long long* sums = a.alloc<unsigned long long>(2);
    fill_i32(err, 0, 1, s);
    fill_i64(reinterpret_cast<int64_t*>(sums), 0, 2, s);
    eoff_t m = read_scalar(rowptr + n, s);
    launch_validate_offsets(n, rowptr, m, err, s);
    if (read_scalar(err, s)) return 1;
    launch_validate(n, m, col, ew, vw, err, sums, s);
    int32_t e = 0;
    unsigned long long h[2];
    JET_HIP(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s));
    JET_HIP(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, s));
    JET_HIP(hipStreamSynchronize(s));
    if (h[0] >= (1ull << 31) || h[1] >= (1ull << 31)) e |= 16;
    return e;
}

__global__ void k_widen_offsets(int64_t n1, const int32_t* in, int64_t* out) { FOR_THREAD(i, n1) out[i] = in[i]; }

void widen_offsets(int64_t n1, const int32_t* in, int64_t* out, hipStream_t s) {
    k_widen_offsets<<<grid_for(n1, 256, 8192), 256, 0, s>>>(n1, in, out);
    JET_LAUNCH_CHECK();
}

// Narrowing with a range check: err |= 32 when a value is outside (0, 2^31).
__global__ void k_narrow(int64_t cnt, const int64_t* in, int32_t* out, int32_t* err) {
    FOR_THREAD(i, cnt) {
        int64_t v = in[i];
        if (v <= 0 || v > 2147483647LL) atomicOr(err, 32);
        out[i] = (int32_t)v;
    }
}

int32_t narrow_weights(int64_t cnt, const int64_t* in, int32_t* out, Arena& a, hipStream_t s) {
    if (cnt <= 0) return 0;
    ArenaScope sc(a);
    int32_t* err = a.alloc<int32_t>(1);
    fill_i32(err, 0, 1, s);
    k_narrow<<<grid_for(cnt, 256, 8192), 256, 0, s>>>(cnt, in, out, err);
    JET_LAUNCH_CHECK();
    return read_scalar(err, s);
}

// Column narrowing with a range check: err |= 2 when a column is outside [0, n).
__global__ void k_narrow_cols(int64_t cnt, const int64_t* in, int64_t n, int32_t* out, int32_t* err) {
    FOR_THREAD(i, cnt) {
        int64_t v = in[i];
        if (v < 0 || v >= n) atomicOr(err, 2);
        out[i] = (int32_t)v;
    }
}

int32_t narrow_columns(int64_t cnt, const int64_t* in, int64_t n, int32_t* out, Arena& a, hipStream_t s) {
    if (cnt <= 0) return 0;
    ArenaScope sc(a);
    int32_t* err = a.alloc<int32_t>(1);
    fill_i32(err, 0, 1, s);
    k_narrow_cols<<<grid_for(cnt, 256, 8192), 256, 0, s>>>(cnt, in, n, out, err);
    JET_LAUNCH_CHECK();
    return read_scalar(err, s);
}

__global__ void k_check_parts(int64_t n, const int32_t* part, int32_t k, int32_t* err) {
    int32_t e = 0;
    FOR_THREAD(i, n) {
        int32_t p = part[i];
        if (p < 0 || p >= k) e = 1;
    }
    if (e) atomicOr(err, 1);
}

bool parts_in_range(int64_t n, const int32_t* part, int32_t k, Arena& a, hipStream_t s) {
    if (n <= 0) return true;
    ArenaScope sc(a);
    int32_t* err = a.alloc<int32_t>(1);
    fill_i32(err, 0, 1, s);
    k_check_parts<<<grid_for(n, 256, 4096), 256, 0, s>>>(n, part, k, err);
    JET_LAUNCH_CHECK();
    return read_scalar(err, s) == 0;
}

__global__ void k_conn_capacity(vid_t n, const eoff_t* rowptr, int32_t k, unsigned long long* out) {
    __shared__ int64_t red[256 / WAVE];
    int64_t c = 0;
    FOR_THREAD(i, n) {
        int64_t d = rowptr[i + 1] - rowptr[i];
        c += d > k ? k : d;
    }
    c = block_reduce_sum<int64_t, 256>(c, red);
    if (threadIdx.x == 0 && c) atomicAdd(out, (unsigned long long)c);
}

int64_t conn_capacity(vid_t n, const eoff_t* rowptr, int32_t k, Arena& a, hipStream_t s) {
    ArenaScope sc(a);
    unsigned long long* out = a.alloc<unsigned long long>(1);
    fill_i64(reinterpret_cast<int64_t*>(out), 0, 1, s);
    if (n > 0) {
        k_conn_capacity<<<grid_for(n, 256, 4096), 256, 0, s>>>(n, rowptr, k, out);
        JET_LAUNCH_CHECK();
    }
    return (int64_t)read_scalar(out, s);
}

__global__ void k_gain_bucket(int64_t cnt, const int64_t* g, const int64_t* w, int32_t* out) {
    FOR_THREAD(i, cnt) out[i] = gain_bucket_fast(g[i], w[i]);
}

void gain_bucket_batch(int64_t cnt, const int64_t* g, const int64_t* w, int32_t* out, hipStream_t s) {
    if (cnt <= 0) return;
    k_gain_bucket<<<grid_for(cnt, 256, 8192), 256, 0, s>>>(cnt, g, w, out);
    JET_LAUNCH_CHECK();
}

}  // namespace jet
