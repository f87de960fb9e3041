// jet_edge_cut_noalloc: the edge cut of jet_edge_cut (part_stat.hpp:77-108)
// with no allocation at all, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// Its only device memory besides the caller's arrays is `scratch`, a
// file-scope array that every GPU the library runs on holds until the process
// ends: the validation error word, the two weight totals of the validation
// and the cut accumulator. None of them grows with the graph (the kernels
// stride over any n and m), so the array is NOALLOC_SCRATCH_BYTES whatever the
// input. Each device's copy of the array sits at its own address, taken once
// per device with hipGetSymbolAddress; host code uses only that address, never
// the array's name. Calls on one device are serialised by a per-device lock
// (they would otherwise race on the shared array, which keeps the last call's
// contents); every call clears the words it accumulates into, with a fill
// kernel (no memset node, so the sequence stays valid under graph capture).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.h"
#include "jet_capi.h"
#include "misc.h"
#include "prims.h"
#include "refine.h"

namespace jet {

constexpr int NOALLOC_WORDS = 32;  // 256 bytes: [0] error bits, [1] edge weight total, [2] vertex weight total, [3] cut
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_WORDS * sizeof(unsigned long long);
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned long long scratch[NOALLOC_WORDS];
static_assert(sizeof(scratch) == NOALLOC_SCRATCH_BYTES, "scratch is the stated 256 bytes");

__global__ void k_noalloc_clear(unsigned long long* w, int count) {
    if ((int)threadIdx.x < count) w[threadIdx.x] = 0;
}

namespace {

struct DeviceScratch {
    std::once_flag once;
    hipError_t rc = hipSuccess;
    unsigned long long* ptr = nullptr;
    std::mutex mu;  // one call at a time per device
};

DeviceScratch g_dev[NOALLOC_MAX_DEVICES];

// The device's copy of `scratch`, resolved with that device current.
unsigned long long* scratch_of(int dev) {
    DeviceScratch& d = g_dev[dev];
    std::call_once(d.once, [&] {
        d.rc = hipGetSymbolAddress(reinterpret_cast<void**>(&d.ptr), HIP_SYMBOL(scratch));
        if (d.rc != hipSuccess) {
            (void)hipGetLastError();
            d.ptr = nullptr;
        }
    });
    if (d.rc != hipSuccess || !d.ptr)
        throw DeviceError(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(d.rc));
    return d.ptr;
}

template <class T>
T read_back(const T* dev_src, hipStream_t s) {
    T v{};
    JET_HIP(hipMemcpyAsync(&v, dev_src, sizeof(T), hipMemcpyDeviceToHost, s));
    JET_HIP(hipStreamSynchronize(s));
    return v;
}

int fail(int code, const std::string& msg) {
    set_last_error(msg);
    return code;
}

}  // namespace

static int edge_cut_noalloc(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                            const int32_t* part, int64_t* cut2_out) {
    if (!row_offsets || !col || !part || !cut2_out) return fail(JET_EINVAL, "row_offsets, col, part and cut2_out are required");
    if (n < 1 || n > (int64_t)ORD_MAX - 2) return fail(JET_EINVAL, "n must be in [1, 2^31 - 2)");
    int dev = 0;
    JET_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(JET_EDEVICE, "device index beyond the supported 64");
    unsigned long long* w = scratch_of(dev);
    std::lock_guard<std::mutex> lk(g_dev[dev].mu);
    const hipStream_t s = nullptr;  // the null stream: nothing to create
    int32_t* err = reinterpret_cast<int32_t*>(w);
    unsigned long long* sums = w + 1;
    unsigned long long* cut = w + 3;
    k_noalloc_clear<<<1, 64, 0, s>>>(w, 4);
    JET_LAUNCH_CHECK();
    // the same checks, in the same order, as the staging of jet_edge_cut
    const int64_t first = read_back(row_offsets, s);
    const int64_t m = read_back(row_offsets + n, s);
    if (first != 0) return fail(JET_EINVAL, "row_offsets[0] must be 0");
    if (m < 0) return fail(JET_EINVAL, "row_offsets[n] must be >= 0");
    launch_validate_offsets((vid_t)n, row_offsets, m, err, s);
    if (read_back(err, s)) return fail(JET_EINVAL, "invalid graph: row offsets not monotone;");
    launch_validate((vid_t)n, m, col, ew, nullptr, err, sums, s);
    int32_t e = read_back(err, s);
    unsigned long long tot[2];
    JET_HIP(hipMemcpyAsync(tot, sums, sizeof(tot), hipMemcpyDeviceToHost, s));
    JET_HIP(hipStreamSynchronize(s));
    if (tot[0] >= (1ull << 31) || tot[1] >= (1ull << 31)) e |= 16;
    if (e) {
        std::string msg = "invalid graph:";
        if (e & 2) msg += " column out of range;";
        if (e & 4) msg += " non-positive edge weight;";
        if (e & 16) msg += " total weight does not fit in int32;";
        return fail(JET_EINVAL, msg);
    }
    DGraph g;
    g.n = (vid_t)n;
    g.m = m;
    g.rowptr = row_offsets;
    g.col = col;
    g.ew = ew;
    launch_cut2(g, part, cut, s);
    *cut2_out = (int64_t)read_back(cut, s);
    set_last_error("");
    return JET_OK;
}

}  // namespace jet

extern "C" int jet_edge_cut_noalloc(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                                    const int32_t* part, int64_t* cut2_out) {
    try {
        return jet::edge_cut_noalloc(n, row_offsets, col, ew, part, cut2_out);
    } catch (const jet::DeviceError& e) {
        jet::set_last_error(e.what());
        return JET_EDEVICE;
    } catch (const std::exception& e) {
        jet::set_last_error(e.what());
        return JET_EDEVICE;
    }
}
