// Connectivity-table operations shared by the refinement kernels
// (jet_refiner.hpp:211-248 reads, :785-886 update_small, :891-901 lookup,
// :1008-1037 shrink rule, :275-307 afterburner edge term). Device only.
#pragma once
#include <hip/hip_runtime.h>

#include "jet_types.h"

namespace jet {

__device__ inline int32_t ld_rx(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int32_t lookup(const int2* tab, int32_t size, part_t target) {
    for (int32_t q = 0; q < size; q++) {
        int32_t pi = (target + q) % size;
        int2 e = tab[pi];
        if (e.x == target) return e.y;
        if (e.x == NULL_PART) return 0;
    }
    return 0;
}

__device__ inline int32_t shrink_size(int32_t used, int32_t cap) {
    int32_t q = used / 4;
    if (q < 3) q = 3;
    int32_t sz = used + q;
    return sz < cap ? sz : cap;
}

// Table slots in groups of four independent loads; f(slot) sees the slots in
// order.
template <class F>
__device__ inline void for_slots(const int2* tab, int32_t sz, F&& f) {
    for (int32_t s0 = 0; s0 < sz; s0 += 4) {
        int2 e[4];
#pragma unroll
        for (int u = 0; u < 4; u++) e[u] = s0 + u < sz ? tab[s0 + u] : make_int2(NULL_PART, 0);
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (s0 + u < sz) f(e[u]);
    }
}

// jet_refiner.hpp:275-307 (afterburner): gain correction assuming every
// higher-priority neighbour (larger pre-gain, or equal and smaller id) moves.
// Non-candidates (destination NO_MOVE) have the lowest priority, as the
// reference's minimum pre-gain gives them.
__device__ inline int64_t afterburn_edge(vid_t v, vid_t i, int32_t ig, int32_t w, part_t p, part_t best,
                                         const part_t* part, const part_t* dest, const int32_t* pregain) {
    // loaded at once; the pre-gain of a non-candidate is stale and unused
    const part_t vd = dest[v], vq = part[v];
    const int32_t vg = pregain[v];
    if (vd == NO_MOVE) return 0;
    int64_t ch = 0;
    if (vg > ig || (vg == ig && v < i)) {
        if (vd == p) ch -= w;
        else if (vd == best) ch += w;
        if (vq == p) ch += w;
        else if (vq == best) ch -= w;
    }
    return ch;
}

__device__ inline int32_t next_slot(int32_t s, int32_t vsz) { return s + 1 == vsz ? 0 : s + 1; }

// update_small subtract (jet_refiner.hpp:785-812). The linear probe for p
// reads four slots per round trip; during the subtraction pass the slot that
// holds p does not move (only parts whose count reaches 0 are reclaimed), so
// the slot found is the one the one-at-a-time probe finds.
__device__ inline void conn_sub(int2* tab, int32_t vsz, part_t p, int32_t w, int32_t k, int32_t* err) {
    int32_t po = p % vsz;
    for (int32_t probed = 0;; probed += 4) {
        if (probed > vsz) {
            atomicOr(err, 1);
            return;
        }
        const int32_t a = po, b = next_slot(a, vsz), c = next_slot(b, vsz), d = next_slot(c, vsz);
        const int32_t xa = ld_rx(&tab[a].x), xb = ld_rx(&tab[b].x), xc = ld_rx(&tab[c].x), xd = ld_rx(&tab[d].x);
        if (xa == p) break;
        if (xb == p) {
            po = b;
            break;
        }
        if (xc == p) {
            po = c;
            break;
        }
        if (xd == p) {
            po = d;
            break;
        }
        po = next_slot(d, vsz);
    }
    int32_t x = atomicAdd(&tab[po].y, -w);
    if (vsz < k && x == w) __hip_atomic_store(&tab[po].x, HASH_RECLAIM, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// update_small add (jet_refiner.hpp:815-886).
__device__ inline void conn_add(int2* tab, int32_t* vsize, int32_t cap, part_t best, int32_t w, int32_t* err) {
    int32_t vsz = ld_rx(vsize);
    int32_t po = best % vsz;
    bool success = false;
    // the search for best or an empty slot, four slots per round trip (the
    // values are those a one-at-a-time probe could read: other threads only
    // fill empty or reclaimed slots meanwhile)
    for (int32_t q = 0, s0 = po; q < vsz; q += 4) {
        int32_t sl[4], e[4];
        sl[0] = s0;
        for (int j = 1; j < 4; j++) sl[j] = next_slot(sl[j - 1], vsz);
        for (int j = 0; j < 4; j++) e[j] = ld_rx(&tab[sl[j]].x);
        bool stop = false;
        for (int j = 0; j < 4 && !stop; j++) {
            if (q + j >= vsz) {
                stop = true;
            } else if (e[j] == best) {
                success = true;
                po = sl[j];
                stop = true;
            } else if (e[j] == NULL_PART) {
                stop = true;
            }
        }
        if (stop) break;
        s0 = next_slot(sl[3], vsz);
    }
    int32_t count = 0;
    while (!success && count < vsz) {
        int32_t px = ld_rx(&tab[po].x);
        while (px != best && px > NULL_PART && count++ < vsz) {
            po = (po + 1) % vsz;
            px = ld_rx(&tab[po].x);
        }
        if (px == best) {
            success = true;
        } else if (px <= NULL_PART) {
            int32_t old = atomicCAS(&tab[po].x, px, best);
            if (old == best || old <= NULL_PART) {
                success = true;
            } else {
                po = (po + 1) % vsz;
                count++;
            }
        }
    }
    if (!success) {
        po = vsz;
        while (!success) {
            if (po >= cap) {
                atomicOr(err, 2);
                return;
            }
            int32_t px = ld_rx(&tab[po].x);
            while (px != best && px > NULL_PART) {
                po++;
                if (po >= cap) {
                    atomicOr(err, 2);
                    return;
                }
                px = ld_rx(&tab[po].x);
            }
            if (px == best) {
                success = true;
            } else {
                int32_t py = atomicCAS(&tab[po].x, px, best);
                if (py == px) {
                    py = best;
                    atomicAdd(vsize, 1);
                }
                if (py == best) success = true;
                else po++;
            }
        }
    }
    atomicAdd(&tab[po].y, w);
}

}  // namespace jet
