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

}  // namespace jet
