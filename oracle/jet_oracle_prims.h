// TEST INFRASTRUCTURE — CPU oracle for the Jet hot path.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
// anything under oracle/. The product never links or calls it.
//
// Scalar primitives shared by every oracle function.
#pragma once
#include <cstdint>
#include <cstddef>
#include <cmath>

namespace orc {

static const int32_t ORD_MAX = 2147483647;
static const int32_t NULL_PART = -1;      // jet_refiner.hpp:91
static const int32_t HASH_RECLAIM = -2;   // jet_refiner.hpp:92
static const int32_t NO_MOVE = -3;        // jet_refiner.hpp:93
static const int32_t MAX_SECTIONS = 128;  // jet_refiner.hpp:95
static const int32_t MAX_BUCKETS = 50;    // jet_refiner.hpp:96
static const int32_t MID_BUCKET = 25;     // jet_refiner.hpp:97

// Random-number sites. The reference draws from Kokkos::Random_XorShift64_Pool
// seeded with std::time (contract.hpp:503); its streams are third-party and
// unpinned. Both the GPU path and this oracle use one counter-based generator
// instead: value = mix64(mix64(seed + site*A + level*B + pass*C) + idx).
enum : uint32_t { SITE_HEC = 1, SITE_MATCH_INIT = 2, SITE_PICK = 3 };

// splitmix64 output function applied to state x (i.e. the first output of a
// splitmix64 generator whose state is x). Vigna's published constants.
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t rng(uint64_t seed, uint32_t level, uint32_t site, uint32_t pass, uint32_t idx) {
    uint64_t s = mix64(seed + (uint64_t)site * 0xD6E8FEB86659FD93ULL
                            + (uint64_t)level * 0xA0761D6478BD642FULL
                            + (uint64_t)pass * 0xE7037ED1A0B428DBULL);
    return mix64(s + (uint64_t)idx);
}

static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// MurmurHash3_x86_32 (Austin Appleby, public domain), general length.
static inline uint32_t murmur3_32(const void* key, size_t len, uint32_t seed) {
    const uint8_t* data = (const uint8_t*)key;
    const size_t nblocks = len / 4;
    uint32_t h1 = seed;
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    for (size_t i = 0; i < nblocks; i++) {
        uint32_t k1 = (uint32_t)data[4 * i] | ((uint32_t)data[4 * i + 1] << 8) |
                      ((uint32_t)data[4 * i + 2] << 16) | ((uint32_t)data[4 * i + 3] << 24);
        k1 *= c1; k1 = rotl32(k1, 15); k1 *= c2;
        h1 ^= k1; h1 = rotl32(h1, 13); h1 = h1 * 5 + 0xe6546b64u;
    }
    const uint8_t* tail = data + nblocks * 4;
    uint32_t k1 = 0;
    switch (len & 3) {
        case 3: k1 ^= (uint32_t)tail[2] << 16; [[fallthrough]];
        case 2: k1 ^= (uint32_t)tail[1] << 8; [[fallthrough]];
        case 1: k1 ^= tail[0]; k1 *= c1; k1 = rotl32(k1, 15); k1 *= c2; h1 ^= k1;
    }
    h1 ^= (uint32_t)len;
    h1 ^= h1 >> 16; h1 *= 0x85ebca6bu; h1 ^= h1 >> 13; h1 *= 0xc2b2ae35u; h1 ^= h1 >> 16;
    return h1;
}

// Kokkos::pod_hash<int32_t> (used at heuristics.hpp:75,474-476,583-586) is
// MurmurHash3_x86_32 over the 4 little-endian bytes of the key with seed 0.
static inline uint32_t pod_hash_i32(int32_t key) { return murmur3_32(&key, 4, 0); }

// heuristics.hpp:49-56, instantiated with uint32_t (pickMatch, :375,414,418).
static inline uint32_t xorshift32(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// jet_refiner.hpp:326-356. The arithmetic is fp32 for the quotient and the
// running value; `gain *= 1.5` / `gain /= 1.5` promote to double and round
// back to float (C++ usual conversions). `abs` resolves to the float overload.
static inline int32_t gain_bucket(int64_t gx, int64_t vwgt) {
    float gain = (float)gx / (float)vwgt;
    int32_t gain_type;
    if (gain > 0.0) {
        gain_type = 0;
    } else if (gain == 0.0) {
        gain_type = 1;
    } else {
        gain_type = MID_BUCKET;
        gain = std::fabs(gain);
        if (gain < 1.0) {
            while (gain < 1.0) {
                gain = (float)((double)gain * 1.5);
                gain_type--;
            }
            if (gain_type < 2) gain_type = 2;
        } else {
            while (gain > 1.0) {
                gain = (float)((double)gain / 1.5);
                gain_type++;
            }
            if (gain_type > MAX_BUCKETS) gain_type = MAX_BUCKETS - 1;
        }
    }
    return gain_type;
}

}  // namespace orc
