// Coarse-graph contraction for gfx950 (reference: src/contract.hpp:347-446).
//
// For a fixed fine->coarse map the coarse adjacency is P^T A P without its
// diagonal, with coarse vertex weights P^T vw. Each coarse row c is built by
// hashing the mapped neighbours of all fine members of c:
//  * rows with at most 128 incident fine edges: one wave per row, open
//    addressing in a per-wave LDS table, ranks by counting;
//  * rows with at most 4096: one workgroup per row, LDS table + bitonic sort;
//  * larger rows (hub aggregates): global spill path, a radix sort of
//    (row, neighbour) keys followed by a reduce-by-key.
// Rows are emitted in increasing neighbour order, which is one of the orders
// the reference's unordered consolidate can produce; weights are integer sums.
#include "contract.h"

#include <algorithm>

#include "tuning.h"
#include "prims.h"
#include "wave.h"

namespace jet {

constexpr int32_t T1_MAX = 128;   // hsize bound for the wave tier
constexpr int32_t T1_CAP = 256;   // wave LDS table slots
constexpr int32_t T2_MAX = 4096;  // hsize bound for the workgroup tier
constexpr int32_t T2_CAP = 8192;  // workgroup LDS table slots
constexpr int32_t EMPTY = -1;

__device__ inline uint32_t slot_hash(vid_t u) { return (uint32_t)u * 0x9E3779B1u; }

__global__ void k_count(DGraph g, const vid_t* map, eoff_t* hsize, wgt_t* cvw, int32_t* mcount) {
    FOR_THREAD(i, g.n) {
        vid_t c = map[i];
        atomicAdd((unsigned long long*)&hsize[c], (unsigned long long)(g.rowptr[i + 1] - g.rowptr[i]));
        atomicAdd(&cvw[c], g.vw[i]);
        atomicAdd(&mcount[c], 1);
    }
}

__global__ void k_members(vid_t n, const vid_t* map, const int32_t* moff, int32_t* mfill, vid_t* members) {
    FOR_THREAD(i, n) {
        vid_t c = map[i];
        members[moff[c] + atomicAdd(&mfill[c], 1)] = (vid_t)i;
    }
}

__global__ void __launch_bounds__(256)
    k_tiers(vid_t nc, const eoff_t* hoff, vid_t* l1, vid_t* l2, vid_t* l3, int32_t* cnt, int32_t t1max,
            int32_t t2max) {
    __shared__ int32_t sh[256 / WAVE + 1];
    FOR_BLOCK_UNIFORM(c, nc) {
        const bool in = c < nc;
        const eoff_t hs = in ? hoff[c + 1] - hoff[c] : 0;
        int32_t a = block_append<256>(&cnt[0], in && hs <= t1max, sh);
        int32_t b = block_append<256>(&cnt[1], in && hs > t1max && hs <= t2max, sh);
        int32_t d = block_append<256>(&cnt[2], in && hs > t2max, sh);
        if (a >= 0) l1[a] = (vid_t)c;
        if (b >= 0) l2[b] = (vid_t)c;
        if (d >= 0) l3[d] = (vid_t)c;
    }
}

// Flattened walk over the edges of a coarse row's fine members, in chunks of
// up to 64 members: lane-strided over edges, member found by binary search on
// the chunk's degree prefix held in LDS.
template <class F>
__device__ inline void wave_row_edges(const DGraph& g, const vid_t* members, int32_t mb, int32_t me, eoff_t* s_ex,
                                      eoff_t* s_rs, F&& f) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int32_t m0 = mb; m0 < me; m0 += WAVE) {
        int32_t mi = m0 + lane;
        bool valid = mi < me;
        eoff_t rs = 0, d = 0;
        if (valid) {
            vid_t i = members[mi];
            rs = g.rowptr[i];
            d = g.rowptr[i + 1] - rs;
        }
        eoff_t inc = wave_incl_scan(d);
        eoff_t total = __shfl(inc, WAVE - 1, WAVE);
        s_ex[lane] = inc - d;
        s_rs[lane] = rs;
        int32_t nm = min(WAVE, me - m0);
        wave_sync();
        for (eoff_t fi = lane; fi < total; fi += WAVE) {
            int lo = 0, hi = nm - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (s_ex[mid] <= fi) lo = mid;
                else hi = mid - 1;
            }
            f(s_rs[lo] + (fi - s_ex[lo]));
        }
        wave_sync();
    }
}

__device__ inline void lds_insert(int32_t* keys, int32_t* vals, uint32_t mask, vid_t u, wgt_t w) {
    uint32_t h = slot_hash(u) & mask;
    while (true) {
        int32_t k = keys[h];
        if (k == u) break;
        if (k == EMPTY) {
            int32_t old = atomicCAS(&keys[h], EMPTY, u);
            if (old == EMPTY || old == u) break;
        }
        h = (h + 1) & mask;
    }
    atomicAdd(&vals[h], w);
}

constexpr int T1_WAVES = 4;

__global__ void __launch_bounds__(T1_WAVES* WAVE)
    k_dedupe_wave(DGraph g, const vid_t* map, const vid_t* members, const int32_t* moff, const eoff_t* hoff,
                  const vid_t* list, const int32_t* cnt, vid_t* tkey, wgt_t* tval, int32_t* ucount) {
    __shared__ int32_t keys[T1_WAVES][T1_CAP];
    __shared__ int32_t vals[T1_WAVES][T1_CAP];
    __shared__ int32_t ukeys[T1_WAVES][T1_CAP];
    __shared__ int32_t uvals[T1_WAVES][T1_CAP];
    __shared__ eoff_t s_ex[T1_WAVES][WAVE];
    __shared__ eoff_t s_rs[T1_WAVES][WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int32_t nrows = cnt[0];
    for (int32_t r = blockIdx.x * T1_WAVES + w; r < nrows; r += gridDim.x * T1_WAVES) {
        vid_t c = list[r];
        eoff_t hs = hoff[c];
        int32_t hsz = (int32_t)(hoff[c + 1] - hs);
        uint32_t cap = 64;
        while ((int32_t)cap < 2 * hsz) cap <<= 1;
        for (uint32_t t = lane; t < cap; t += WAVE) {
            keys[w][t] = EMPTY;
            vals[w][t] = 0;
        }
        wave_sync();
        int32_t* K = keys[w];
        int32_t* V = vals[w];
        wave_row_edges(g, members, moff[c], moff[c + 1], s_ex[w], s_rs[w], [&](eoff_t j) {
            vid_t u = map[g.col[j]];
            if (u != c) lds_insert(K, V, cap - 1, u, g.ew[j]);
        });
        wave_sync();
        // compact unique entries (slot order), then rank by counting
        int32_t base = 0;
        for (uint32_t t0 = 0; t0 < cap; t0 += WAVE) {
            uint32_t t = t0 + lane;
            int32_t k = t < cap ? K[t] : EMPTY;
            bool has = k != EMPTY;
            uint64_t bal = __ballot(has);
            int32_t pos = base + (int32_t)__popcll(bal & ((1ull << lane) - 1ull));
            if (has) {
                ukeys[w][pos] = k;
                uvals[w][pos] = V[t];
            }
            base += (int32_t)__popcll(bal);
        }
        wave_sync();
        const int32_t U = base;
        for (int32_t e = lane; e < U; e += WAVE) {
            int32_t k = ukeys[w][e];
            int32_t rank = 0;
            for (int32_t q = 0; q < U; q++) rank += ukeys[w][q] < k ? 1 : 0;
            tkey[hs + rank] = k;
            tval[hs + rank] = uvals[w][e];
        }
        if (lane == 0) ucount[c] = U;
        wave_sync();
    }
}

constexpr int T2_BLOCK = 256;

__global__ void __launch_bounds__(T2_BLOCK)
    k_dedupe_block(DGraph g, const vid_t* map, const vid_t* members, const int32_t* moff, const eoff_t* hoff,
                   const vid_t* list, const int32_t* cnt, vid_t* tkey, wgt_t* tval, int32_t* ucount) {
    __shared__ int32_t keys[T2_CAP];
    __shared__ int32_t vals[T2_CAP];
    __shared__ int32_t ucnt;
    const int32_t nrows = cnt[1];
    for (int32_t r = blockIdx.x; r < nrows; r += gridDim.x) {
        vid_t c = list[r];
        eoff_t hs = hoff[c];
        int32_t hsz = (int32_t)(hoff[c + 1] - hs);
        uint32_t cap = 256;
        while ((int32_t)cap < 2 * hsz) cap <<= 1;
        for (uint32_t t = threadIdx.x; t < cap; t += T2_BLOCK) {
            keys[t] = EMPTY;
            vals[t] = 0;
        }
        if (threadIdx.x == 0) ucnt = 0;
        __syncthreads();
        int32_t mb = moff[c], me = moff[c + 1];
        const int lane = threadIdx.x & (WAVE - 1);
        const int wid = threadIdx.x / WAVE;
        // members round-robin over waves; each wave walks its member's row
        for (int32_t mi = mb + wid; mi < me; mi += T2_BLOCK / WAVE) {
            vid_t i = members[mi];
            for (eoff_t j = g.rowptr[i] + lane; j < g.rowptr[i + 1]; j += WAVE) {
                vid_t u = map[g.col[j]];
                if (u != c) lds_insert(keys, vals, cap - 1, u, g.ew[j]);
            }
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < cap; t += T2_BLOCK)
            if (keys[t] != EMPTY) atomicAdd(&ucnt, 1);
        // bitonic sort of (key as unsigned, EMPTY last) over cap slots
        for (uint32_t size = 2; size <= cap; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (uint32_t t = threadIdx.x; t < cap / 2; t += T2_BLOCK) {
                    uint32_t lo = 2 * t - (t & (stride - 1));
                    uint32_t hi = lo + stride;
                    bool up = (lo & size) == 0;
                    uint32_t a = (uint32_t)keys[lo], b = (uint32_t)keys[hi];
                    if ((a > b) == up) {
                        keys[lo] = (int32_t)b;
                        keys[hi] = (int32_t)a;
                        int32_t tv = vals[lo];
                        vals[lo] = vals[hi];
                        vals[hi] = tv;
                    }
                }
            }
        }
        __syncthreads();
        const int32_t U = ucnt;
        for (int32_t e = threadIdx.x; e < U; e += T2_BLOCK) {
            tkey[hs + e] = keys[e];
            tval[hs + e] = vals[e];
        }
        if (threadIdx.x == 0) ucount[c] = U;
        __syncthreads();
    }
}

// ---- spill path: (row rank << 32 | neighbour) keys, radix sort, reduce by key
__global__ void __launch_bounds__(256)
    k_spill_gather(DGraph g, const vid_t* map, const vid_t* members, const int32_t* moff, const vid_t* list,
                   int32_t n3, const eoff_t* soff, uint64_t* skey, uint32_t* sval) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    __shared__ eoff_t wpos[256 / WAVE];
    for (int32_t r = blockIdx.x; r < n3; r += gridDim.x) {
        vid_t c = list[r];
        eoff_t base = soff[r];
        int32_t mb = moff[c], me = moff[c + 1];
        // members round-robin over waves; wave-private output cursors derived
        // from the member's position in the row's member prefix
        for (int32_t mi0 = mb; mi0 < me; mi0 += 256 / WAVE) {
            int32_t mi = mi0 + wid;
            eoff_t d = 0, rs = 0;
            if (mi < me) {
                vid_t i = members[mi];
                rs = g.rowptr[i];
                d = g.rowptr[i + 1] - rs;
            }
            if (lane == 0) wpos[wid] = d;
            __syncthreads();
            eoff_t before = 0;
            for (int q = 0; q < wid; q++) before += wpos[q];
            eoff_t chunk = 0;
            for (int q = 0; q < 256 / WAVE; q++) chunk += wpos[q];
            for (eoff_t t = lane; t < d; t += WAVE) {
                vid_t u = map[g.col[rs + t]];
                uint64_t key = u == c ? ~0ull : (((uint64_t)r << 32) | (uint32_t)u);
                skey[base + before + t] = key;
                sval[base + before + t] = (uint32_t)g.ew[rs + t];
            }
            __syncthreads();
            base += chunk;
        }
    }
}

__global__ void k_spill_heads(const uint64_t* skey, int64_t len, int32_t* head) {
    FOR_THREAD(t, len) {
        uint64_t k = skey[t];
        head[t] = (k != ~0ull && (t == 0 || k != skey[t - 1])) ? 1 : 0;
    }
}

__global__ void k_spill_reduce(const uint64_t* skey, const uint32_t* sval, int64_t len, const int32_t* uid_ex,
                               const vid_t* list, int32_t* ukey, wgt_t* uval, int32_t* urow, int32_t* rowfirst) {
    FOR_THREAD(t, len) {
        uint64_t k = skey[t];
        if (k == ~0ull) continue;
        int32_t uid = uid_ex[t + 1] - 1;
        atomicAdd(&uval[uid], (wgt_t)sval[t]);
        if (uid_ex[t + 1] != uid_ex[t]) {
            int32_t r = (int32_t)(k >> 32);
            ukey[uid] = (int32_t)(k & 0xffffffffull);
            urow[uid] = r;
            atomicMin(&rowfirst[r], uid);
        }
    }
}

__global__ void k_spill_write(int32_t U, const int32_t* ukey, const wgt_t* uval, const int32_t* urow,
                              const int32_t* rowfirst, const vid_t* list, const eoff_t* hoff, vid_t* tkey, wgt_t* tval,
                              int32_t* ucount) {
    FOR_THREAD(e, U) {
        int32_t r = urow[e];
        vid_t c = list[r];
        int32_t off = (int32_t)e - rowfirst[r];
        tkey[hoff[c] + off] = ukey[e];
        tval[hoff[c] + off] = uval[e];
        atomicAdd(&ucount[c], 1);
    }
}

__global__ void k_consolidate(vid_t nc, int64_t mc, const eoff_t* crow, const eoff_t* hoff, const vid_t* tkey,
                              const wgt_t* tval, vid_t* col, wgt_t* ew) {
    FOR_THREAD(e, mc) {
        int32_t lo = 0, hi = nc - 1;
        while (lo < hi) {
            int32_t mid = (lo + hi + 1) >> 1;
            if (crow[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        eoff_t src = hoff[lo] + (e - crow[lo]);
        col[e] = tkey[src];
        ew[e] = tval[src];
    }
}

__global__ void k_spill_sizes(const vid_t* list, int32_t n3, const eoff_t* hoff, eoff_t* ssz) {
    FOR_THREAD(r, n3) {
        vid_t c = list[r];
        ssz[r] = hoff[c + 1] - hoff[c];
    }
}

void contract(const DGraph& g, const vid_t* map, vid_t nc, Arena& persist, Arena& a, hipStream_t s, Level& out) {
    out.n = nc;
    out.rowptr = persist.alloc<eoff_t>((int64_t)nc + 1);
    out.vw = persist.alloc<wgt_t>(nc);
    ArenaScope sc(a);
    eoff_t* hoff = a.alloc<eoff_t>((int64_t)nc + 1);
    int32_t* moff = a.alloc<int32_t>((int64_t)nc + 1);
    int32_t* mfill = a.alloc<int32_t>(nc);
    int32_t* ucount = a.alloc<int32_t>((int64_t)nc + 1);
    vid_t* members = a.alloc<vid_t>(g.n);
    vid_t* lists = a.alloc<vid_t>(3 * (int64_t)nc);
    int32_t* cnt = a.alloc<int32_t>(4);
    fill_i64(hoff, 0, (int64_t)nc + 1, s);
    fill_i32(out.vw, 0, nc, s);
    fill_i32(moff, 0, (int64_t)nc + 1, s);
    fill_i32(mfill, 0, nc, s);
    fill_i32(ucount, 0, (int64_t)nc + 1, s);
    fill_i32(cnt, 0, 4, s);
    int nb = grid_for(g.n, 256, 8192);
    k_count<<<nb, 256, 0, s>>>(g, map, hoff, out.vw, moff);
    JET_LAUNCH_CHECK();
    excl_scan_i64_i64(hoff, hoff, nc, a, s);
    excl_scan_i32_i32(moff, moff, nc, a, s);
    k_members<<<nb, 256, 0, s>>>(g.n, map, moff, mfill, members);
    vid_t* l1 = lists;
    vid_t* l2 = lists + nc;
    vid_t* l3 = lists + 2 * (int64_t)nc;
    // the tuned tier bounds never exceed T1_MAX / T2_MAX (half of each table)
    const int32_t t1max = (int32_t)std::min<int64_t>(T1_MAX, tuning().ct_wave_max);
    const int32_t t2max = (int32_t)std::min<int64_t>(T2_MAX, std::max<int64_t>(t1max, tuning().ct_block_max));
    k_tiers<<<grid_for(nc, 256, 8192), 256, 0, s>>>(nc, hoff, l1, l2, l3, cnt, t1max, t2max);
    JET_LAUNCH_CHECK();
    int32_t hcnt[4];
    read_i32_words(cnt, 4, hcnt, s);
    eoff_t H = g.m;
    vid_t* tkey = a.alloc<vid_t>(H);
    wgt_t* tval = a.alloc<wgt_t>(H);
    if (hcnt[0] > 0) {
        k_dedupe_wave<<<grid_for(hcnt[0], T1_WAVES, 8192), T1_WAVES * WAVE, 0, s>>>(g, map, members, moff, hoff, l1,
                                                                                   cnt, tkey, tval, ucount);
        JET_LAUNCH_CHECK();
    }
    if (hcnt[1] > 0) {
        k_dedupe_block<<<std::min(hcnt[1], 4096), T2_BLOCK, 0, s>>>(g, map, members, moff, hoff, l2, cnt, tkey, tval,
                                                                    ucount);
        JET_LAUNCH_CHECK();
    }
    if (hcnt[2] > 0) {
        int32_t n3 = hcnt[2];
        ArenaScope sc3(a);
        eoff_t* soff = a.alloc<eoff_t>((int64_t)n3 + 1);
        k_spill_sizes<<<grid_for(n3, 256, 1024), 256, 0, s>>>(l3, n3, hoff, soff);
        JET_LAUNCH_CHECK();
        excl_scan_i64_i64(soff, soff, n3, a, s);
        eoff_t H3 = read_scalar(soff + n3, s);
        uint64_t* skey = a.alloc<uint64_t>(H3);
        uint32_t* sval = a.alloc<uint32_t>(H3);
        uint64_t* ktmp = a.alloc<uint64_t>(H3);
        uint32_t* vtmp = a.alloc<uint32_t>(H3);
        k_spill_gather<<<std::min(n3, 4096), 256, 0, s>>>(g, map, members, moff, l3, n3, soff, skey, sval);
        JET_LAUNCH_CHECK();
        int rbits = 1;
        while ((int64_t(1) << rbits) < n3 + 1) rbits++;
        radix_sort_u64(skey, sval, H3, 32 + rbits + 1, ktmp, vtmp, a, s);
        int32_t* uid = a.alloc<int32_t>(H3 + 1);
        int hb = grid_for(H3, 256, 8192);
        k_spill_heads<<<hb, 256, 0, s>>>(skey, H3, uid);
        JET_LAUNCH_CHECK();
        excl_scan_i32_i32(uid, uid, H3, a, s);
        int32_t U = read_scalar(uid + H3, s);
        int32_t* ukey = a.alloc<int32_t>(U);
        wgt_t* uval = a.alloc<wgt_t>(U);
        int32_t* urow = a.alloc<int32_t>(U);
        int32_t* rowfirst = a.alloc<int32_t>(n3);
        fill_i32(uval, 0, U, s);
        fill_i32(rowfirst, ORD_MAX, n3, s);
        k_spill_reduce<<<hb, 256, 0, s>>>(skey, sval, H3, uid, l3, ukey, uval, urow, rowfirst);
        JET_LAUNCH_CHECK();
        if (U > 0) {
            k_spill_write<<<grid_for(U, 256, 8192), 256, 0, s>>>(U, ukey, uval, urow, rowfirst, l3, hoff, tkey, tval,
                                                                 ucount);
            JET_LAUNCH_CHECK();
        }
    }
    excl_scan_i32_i64(ucount, out.rowptr, nc, a, s);
    out.m = read_scalar(out.rowptr + nc, s);
    out.col = persist.alloc<vid_t>(out.m);
    out.ew = persist.alloc<wgt_t>(out.m);
    if (out.m > 0) {
        k_consolidate<<<grid_for(out.m, 256, 16384), 256, 0, s>>>(nc, out.m, out.rowptr, hoff, tkey, tval, out.col,
                                                                   out.ew);
        JET_LAUNCH_CHECK();
    }
}

}  // namespace jet
