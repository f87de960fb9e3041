// Shared types, constants and device helpers of the MI355X Jet hot path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace jet {

using vid_t = int32_t;   // vertex ordinal (ordinal_t in header/jet_defs.h:51)
using eoff_t = int64_t;  // edge offset (big_offset_t in header/jet_defs.h:48)
using wgt_t = int32_t;   // vertex / edge weight and connectivity
using part_t = int32_t;  // part id (part_t in header/jet_defs.h:72)

constexpr int32_t ORD_MAX = 2147483647;
constexpr part_t NULL_PART = -1;     // jet_refiner.hpp:91
constexpr part_t HASH_RECLAIM = -2;  // jet_refiner.hpp:92
constexpr part_t NO_MOVE = -3;       // jet_refiner.hpp:93
constexpr int32_t MAX_SECTIONS = 128;
constexpr int32_t MAX_BUCKETS = 50;
constexpr int32_t MID_BUCKET = 25;
constexpr int32_t GAIN_MIN = -2147483647 - 1;

constexpr int WAVE = 64;

enum : uint32_t { SITE_HEC = 1, SITE_MATCH_INIT = 2, SITE_PICK = 3 };

struct DeviceError : std::runtime_error {
    explicit DeviceError(const std::string& s) : std::runtime_error(s) {}
};

#define JET_HIP(call)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            throw ::jet::DeviceError(std::string(#call) + ": " + hipGetErrorString(e_) + " at " __FILE__); \
    } while (0)

#define JET_LAUNCH_CHECK() JET_HIP(hipGetLastError())

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Device view of one level's graph: CSR with both directions stored.
struct DGraph {
    vid_t n = 0;
    eoff_t m = 0;
    const eoff_t* rowptr = nullptr;
    const vid_t* col = nullptr;
    const wgt_t* ew = nullptr;
    const wgt_t* vw = nullptr;
};

// ---------------------------------------------------------------- hashes
// splitmix64 output function of state z (first output of a generator whose
// state is z); Vigna's constants.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Counter-based generator used in place of the reference's time-seeded
// Kokkos::Random_XorShift64_Pool (contract.hpp:503): one value per
// (seed, level, call site, pass, index).
__host__ __device__ inline uint64_t rng(uint64_t seed, uint32_t level, uint32_t site, uint32_t pass, uint32_t idx) {
    uint64_t s = mix64(seed + (uint64_t)site * 0xD6E8FEB86659FD93ULL + (uint64_t)level * 0xA0761D6478BD642FULL +
                       (uint64_t)pass * 0xE7037ED1A0B428DBULL);
    return mix64(s + (uint64_t)idx);
}

__host__ __device__ inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// Kokkos::pod_hash<int32_t>: MurmurHash3_x86_32 of the key's 4 bytes, seed 0
// (one block, no tail).
__host__ __device__ inline uint32_t pod_hash_i32(int32_t key) {
    uint32_t k1 = (uint32_t)key;
    k1 *= 0xcc9e2d51u;
    k1 = rotl32(k1, 15);
    k1 *= 0x1b873593u;
    uint32_t h1 = k1;
    h1 = rotl32(h1, 13);
    h1 = h1 * 5 + 0xe6546b64u;
    h1 ^= 4u;
    h1 ^= h1 >> 16;
    h1 *= 0x85ebca6bu;
    h1 ^= h1 >> 13;
    h1 *= 0xc2b2ae35u;
    h1 ^= h1 >> 16;
    return h1;
}

// heuristics.hpp:49-56 instantiated with uint32_t.
__host__ __device__ inline uint32_t xorshift32(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// jet_refiner.hpp:326-356. fp32 quotient; each *=1.5 / /=1.5 step is done in
// double and rounded to float, as C++ evaluates `float op double`. The
// quotient is formed in double from the two float operands and rounded once,
// which equals the correctly rounded fp32 division.
__host__ __device__ inline int32_t gain_bucket(int64_t gx, int64_t vwgt) {
    float a = (float)gx;
    float b = (float)vwgt;
    float gain = (float)((double)a / (double)b);
    int32_t gain_type;
    if (gain > 0.0f) {
        gain_type = 0;
    } else if (gain == 0.0f) {
        gain_type = 1;
    } else {
        gain_type = MID_BUCKET;
        gain = -gain;
        if (gain < 1.0f) {
            while (gain < 1.0f) {
                gain = (float)((double)gain * 1.5);
                gain_type--;
            }
            if (gain_type < 2) gain_type = 2;
        } else {
            while (gain > 1.0f) {
                gain = (float)((double)gain / 1.5);
                gain_type++;
            }
            if (gain_type > MAX_BUCKETS) gain_type = MAX_BUCKETS - 1;
        }
    }
    return gain_type;
}

// gain_bucket without the loops. x -> (float)(x / 1.5) (and x * 1.5) is
// monotone, so the number of loop steps from g is the number of thresholds g
// passes: GB_UP[j] is the largest float that reaches <= 1 within j divisions,
// GB_DOWN[j] the smallest that reaches >= 1 within j multiplications. Only the
// first 26 / 23 steps matter: more are clamped. The quotient skips the double
// division when the weight is 1 (the quotient is then the gain itself).
constexpr float GB_UP[26] = {
    0x1.0000000000000p+0f, 0x1.8000000000000p+0f, 0x1.2000000000000p+1f, 0x1.b000000000000p+1f,
    0x1.4400000000000p+2f, 0x1.e600000000000p+2f, 0x1.6c80000000000p+3f, 0x1.1160000000000p+4f,
    0x1.9a10000000000p+4f, 0x1.338c000000000p+5f, 0x1.cd52000000000p+5f, 0x1.59fd800000000p+6f,
    0x1.037e200000000p+7f, 0x1.853d300000000p+7f, 0x1.23ede40000000p+8f, 0x1.b5e4d60000000p+8f,
    0x1.486ba00000000p+9f, 0x1.eca1700000000p+9f, 0x1.7179140000000p+10f, 0x1.151ace0000000p+11f,
    0x1.9fa8360000000p+11f, 0x1.37be280000000p+12f, 0x1.d39d3c0000000p+12f, 0x1.5eb5ec0000000p+13f,
    0x1.0708700000000p+14f, 0x1.8a8ca80000000p+14f};
constexpr float GB_DOWN[23] = {
    0x1.0000000000000p+0f, 0x1.5555560000000p-1f, 0x1.c71c720000000p-2f, 0x1.2f684c0000000p-2f,
    0x1.948b100000000p-3f, 0x1.0db20a0000000p-3f, 0x1.67980e0000000p-4f, 0x1.df75680000000p-5f,
    0x1.3fa39a0000000p-5f, 0x1.aa2f780000000p-6f, 0x1.1c1fa60000000p-6f, 0x1.7ad4de0000000p-7f,
    0x1.f91bd20000000p-8f, 0x1.50bd380000000p-8f, 0x1.c0fc4a0000000p-9f, 0x1.2b52dc0000000p-9f,
    0x1.8f19240000000p-10f, 0x1.0a10c20000000p-10f, 0x1.62c1020000000p-11f, 0x1.d901580000000p-12f,
    0x1.3b563a0000000p-12f, 0x1.a472f80000000p-13f, 0x1.184ca60000000p-13f};

__host__ __device__ inline int32_t gain_bucket_fast(int64_t gx, int64_t vwgt) {
    const float a = (float)gx;
    const float b = (float)vwgt;
    float gain = b == 1.0f ? a : (float)((double)a / (double)b);
    if (gain > 0.0f) return 0;
    if (gain == 0.0f) return 1;
    gain = -gain;
    int32_t c = 0;
    if (gain < 1.0f) {
#pragma unroll
        for (int j = 0; j < 23; j++) c += gain < GB_DOWN[j] ? 1 : 0;
        return MID_BUCKET - c < 2 ? 2 : MID_BUCKET - c;
    }
#pragma unroll
    for (int j = 0; j < 26; j++) c += gain > GB_UP[j] ? 1 : 0;
    return MID_BUCKET + c > MAX_BUCKETS ? MAX_BUCKETS - 1 : MID_BUCKET + c;
}

}  // namespace jet
