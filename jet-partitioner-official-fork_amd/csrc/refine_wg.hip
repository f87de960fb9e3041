// Whole-level Jet refinement in one workgroup of 1024 threads, for small
// levels (reference: jet_refine, src/jet_refiner.hpp:1050-1142, and every
// step it calls).
//
// On a level of a few thousand vertices each step of the multi-kernel path is
// a chain of dependent loads that finishes in a few microseconds, so an
// iteration of about ten launches and one host round trip costs 50-70 us for
// microseconds of work. Here one workgroup runs the iteration loop itself:
// the controller (step choice, best-state rule, counters, temperatures) lives
// in LDS, the steps are separated by workgroup barriers, and the level's
// tables and arrays stay in the CU's caches and the L2. The steps compute what
// the multi-kernel path computes, with the same canonical schedule:
//
// * jet_lp (:211-322): cached destinations, the filter, candidates, the
//   afterburner; ties between parts of equal connectivity go to the lowest id.
// * rebalance_weak / rebalance_strong (:417-645): minibucket sums in LDS, their
//   scan, whole buckets decided by their offsets, and members of the buckets
//   that need positions ranked by (bucket, vertex id) in chunks of 1024
//   (bitonic sort), which is the position order of a serial atomic_fetch_add;
//   the strong destinations with the cookie cutter and the re-draws of the
//   move list in increasing id by one thread.
// * perform_moves (:905-955): part sizes in LDS, the cut change from the parts
//   before and after the moves (equal to the reference's table lookups before
//   and after the updates), update_small's subtractions then additions with
//   the shared table operations, or, for more than n/10 moves, the rows to
//   rebuild marked; the workgroup then stops after the iteration and the host
//   rebuilds them with the multi-workgroup build before relaunching.
//
// Items of degree > P.big_deg (afterburner candidates and moves) are taken by
// a whole wave instead of one thread.
#include "refine_wg.h"

#include <algorithm>
#include <mutex>

#include "conn_table.h"
#include "wave.h"

namespace jet {

constexpr int WG_BLOCK = 1024;
constexpr int WG_WAVES = WG_BLOCK / WAVE;
constexpr int32_t WG_ITER_CAP = 1 << 20;

namespace {
thread_local WgCollector* t_collector = nullptr;
}

void set_wg_collector(WgCollector* c) { t_collector = c; }
WgCollector* wg_collector() { return t_collector; }

bool wg_fits(int64_t n, int32_t k, int64_t nbk, int64_t max_n) {
    return n >= 1 && n <= max_n && k >= 1 && k <= WG_MAX_K && nbk <= WG_MAX_NBK;
}

namespace {

// LDS of the kernel besides the dynamic bucket array.
struct WgShared {
    int64_t sizes[WG_MAX_K];
    int64_t evict_start[WG_MAX_K + 1];
    int64_t evict_adjust[WG_MAX_K];
    part_t undersized[WG_MAX_K];
    uint64_t key[WG_BLOCK];
    int64_t pre[WG_BLOCK];
    int64_t red[WG_WAVES + 1];
    int32_t cnt[8];  // [0] small candidates, [1] big candidates, [2] small moves, [3] big moves, [4] ranked, [5] re-draws
    int32_t nunder;
    int32_t take;
    int32_t rebuild;
    int64_t counter;  // the strong re-draw counter (max_vwgt after the cookie cutter)
    WgState S;
};

__device__ inline int64_t deg_of(const DGraph& g, vid_t i) { return g.rowptr[i + 1] - g.rowptr[i]; }

// Appends v to the front (small) or back (big) of list; counters in LDS.
__device__ inline void list_push(vid_t* list, vid_t n, int32_t* cnt_small, int32_t* cnt_big, bool pred, bool big,
                                 vid_t v) {
    const int32_t a = wave_append(cnt_small, pred && !big);
    const int32_t b = wave_append(cnt_big, pred && big);
    if (a >= 0) list[a] = v;
    if (b >= 0) list[n - 1 - b] = v;
}

// Sum of f over the edges of the listed items: small items (front, ns of
// them) G lanes each (a power of two up to a wave: the lanes of a group split
// the item's edges, so its loads are in flight together), big items (back,
// nb of them) one wave each; item(i, sub, lanes) returns the lane's partial
// sum. Returns each thread's share.
template <class F>
__device__ inline int64_t over_items(const vid_t* list, vid_t n, int32_t ns, int32_t nb, int G, F&& item) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = threadIdx.x / WAVE;
    const int gsub = threadIdx.x & (G - 1);
    const int32_t gid = threadIdx.x / G, ngroups = WG_BLOCK / G;
    int64_t acc = 0;
    for (int32_t x = gid; x < ns; x += ngroups) acc += item(list[x], gsub, G);
    for (int32_t x = wv; x < nb; x += WG_WAVES) acc += item(list[n - 1 - x], lane, WAVE);
    return acc;
}

// Sum of v over the G lanes of each aligned group (G a power of two up to a
// wave), in every lane of the group.
__device__ inline int64_t group_sum(int64_t v, int G) {
    for (int o = 1; o < G; o <<= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)(uint64_t)v, o, WAVE);
        const uint32_t hi = __shfl_xor((uint32_t)((uint64_t)v >> 32), o, WAVE);
        v += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    return v;
}

// Exclusive scan of a[0..len) in place (LDS), whole workgroup; returns the total.
__device__ inline int64_t block_scan_inplace(int32_t* a, int32_t len, int64_t* red) {
    const int32_t per = (len + WG_BLOCK - 1) / WG_BLOCK;
    const int32_t t0 = threadIdx.x * per;
    int64_t sum = 0;
    for (int32_t j = 0; j < per && t0 + j < len; j++) sum += a[t0 + j];
    int64_t tot;
    int64_t run = block_excl_scan<int64_t, WG_BLOCK>(sum, tot, red);
    for (int32_t j = 0; j < per && t0 + j < len; j++) {
        const int32_t v = a[t0 + j];
        a[t0 + j] = (int32_t)run;
        run += v;
    }
    __syncthreads();
    return tot;
}

// Appends the vertices of list[0..len) with pred(v) set to out[base..] in
// list order (each thread takes a contiguous range, one workgroup scan).
// Returns the count.
template <class P>
__device__ inline int32_t ordered_compact(const vid_t* in, int32_t len, vid_t* out, P&& pred, int64_t* red) {
    const int32_t per = (len + WG_BLOCK - 1) / WG_BLOCK;
    const int32_t t0 = threadIdx.x * per;
    int64_t c = 0;
    for (int32_t j = 0; j < per && t0 + j < len; j++) c += pred(in[t0 + j]) ? 1 : 0;
    int64_t tot;
    int64_t at = block_excl_scan<int64_t, WG_BLOCK>(c, tot, red);
    for (int32_t j = 0; j < per && t0 + j < len; j++) {
        const vid_t v = in[t0 + j];
        if (pred(v)) out[at++] = v;
    }
    __syncthreads();
    return (int32_t)tot;
}

// The same over vertex ids 0..n-1.
template <class P>
__device__ inline int32_t ordered_compact_ids(vid_t n, vid_t* out, P&& pred, int64_t* red) {
    const int32_t per = (n + WG_BLOCK - 1) / WG_BLOCK;
    const int32_t t0 = threadIdx.x * per;
    int64_t c = 0;
    for (int32_t j = 0; j < per && t0 + j < n; j++) c += pred(t0 + j) ? 1 : 0;
    int64_t tot;
    int64_t at = block_excl_scan<int64_t, WG_BLOCK>(c, tot, red);
    for (int32_t j = 0; j < per && t0 + j < n; j++)
        if (pred(t0 + j)) out[at++] = t0 + j;
    __syncthreads();
    return (int32_t)tot;
}

// Bitonic sort of one key and weight per thread (ascending keys), exchanges
// within a wave by shuffles, across waves through LDS.
__device__ inline void block_sort(uint64_t& key, uint32_t& w, uint64_t* skey, int64_t* spay) {
    const uint32_t t = threadIdx.x;
    for (uint32_t size = 2; size <= WG_BLOCK; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            uint64_t other;
            uint32_t ow;
            if (stride >= WAVE) {
                skey[t] = key;
                spay[t] = w;
                __syncthreads();
                other = skey[t ^ stride];
                ow = (uint32_t)spay[t ^ stride];
                __syncthreads();
            } else {
                const uint32_t lo = __shfl_xor((uint32_t)key, (int)stride, WAVE);
                const uint32_t hi = __shfl_xor((uint32_t)(key >> 32), (int)stride, WAVE);
                other = ((uint64_t)hi << 32) | lo;
                ow = __shfl_xor(w, (int)stride, WAVE);
            }
            const bool lower = (t & stride) == 0, up = (t & size) == 0;
            const bool take = (lower == up) ? (other < key) : (other > key);
            if (take) {
                key = other;
                w = ow;
            }
        }
    }
}

// First index in [lo, hi) whose evict_start exceeds s (hi when none): the
// reference's forward walk `while (p <= k && evict_start(p) <= s) p++` from lo.
__device__ inline int32_t walk_chunks(const int64_t* es, int32_t lo, int32_t hi, int64_t s) {
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (es[mid] <= s) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

}  // namespace

// The level loop of one workgroup over the level P describes; the final
// controller state goes to *P.st and, when mirror is set, to *mirror.
__device__ __forceinline__ void refine_wg_body(const WgParams& P, WgState* mirror) {
    extern __shared__ int32_t sbk[];  // minibucket sums, then their scan
    __shared__ WgShared L;
    const DGraph g = P.g;
    const vid_t n = g.n;
    const int32_t k = P.k;
    const int t = threadIdx.x;
    const int lane = t & (WAVE - 1);
    const int wv = t / WAVE;
    int32_t* err = &P.st->err;

    if (t == 0) {
        L.S = *P.st;
        L.S.status = WG_DONE;
    }
    for (int32_t p = t; p < k; p += WG_BLOCK) L.sizes[p] = P.curr_sizes[p];
    __syncthreads();

    for (;;) {
        // the temperature loop of :1095-1134: `for (fr : temps) { count = 0;
        // while (count++ <= 11) ... }` as the state (temp_idx, count)
        const int32_t ti = L.S.temp_idx;
        if (ti >= P.ntemps) break;
        if (L.S.iterations >= WG_ITER_CAP) {
            // the controller ends every level after a bounded number of
            // iterations (each reset of `count` needs a strictly smaller
            // imbalance or cut); reaching this means a broken invariant
            __syncthreads();
            if (t == 0) L.S.status = WG_STALLED;
            break;
        }
        const bool next_temp = L.S.count > 11;
        const int64_t curr_imb = L.S.curr_imb;
        const int32_t balance_counter = L.S.balance_counter;
        __syncthreads();
        if (next_temp) {
            if (t == 0) {
                L.S.temp_idx++;
                L.S.count = 0;
            }
            __syncthreads();
            continue;
        }
        const int kind = curr_imb <= P.imb_max ? 0 : (balance_counter < 2 ? 1 : 2);
        if (t < 8) L.cnt[t] = 0;
        if (t == 0) {
            L.S.count++;
            L.S.iterations++;
            L.rebuild = 0;
        }
        __syncthreads();

        if (kind == 0) {
            // ---------------------------------------------------- jet_lp (:211-268)
            const double fr = P.temps[ti];
            for (vid_t i = t; i < n; i += WG_BLOCK) {
                const part_t cached = P.dest_cache[i];
                const int32_t lk = P.lock[i];
                part_t best = cached;
                int32_t gain = 0;
                if (cached == NULL_PART) {
                    const part_t p = P.part[i];
                    best = NO_MOVE;
                    int32_t b_conn = 0, p_conn = 0;
                    for_slots(P.conn + P.conn_off[i], P.conn_size[i], [&](int2 e) {
                        if (e.y > b_conn && e.x != p) {
                            best = e.x;
                            b_conn = e.y;
                        } else if (e.y == b_conn && b_conn > 0 && e.x != p && e.x < best) {
                            best = e.x;
                        } else if (e.y > 0 && e.x == p) {
                            p_conn = e.y;
                        }
                    });
                    if (best != NO_MOVE) {
                        if (b_conn >= p_conn || ((double)(p_conn - b_conn) < floor(fr * (double)p_conn)))
                            gain = b_conn - p_conn;
                        else best = NO_MOVE;
                    }
                    P.gain_p[i] = gain;
                    P.dest_cache[i] = best;
                } else if (cached != NO_MOVE && lk == 0) {
                    gain = P.gain_p[i];
                }
                const bool cand = best != NO_MOVE && lk == 0;
                P.dest[i] = cand ? best : NO_MOVE;
                if (!cand && lk) P.lock[i] = 0;
                bool big = false;
                if (cand) {
                    P.pregain[i] = gain;
                    big = deg_of(g, i) > P.big_deg;
                }
                list_push(P.cand, n, &L.cnt[0], &L.cnt[1], cand, big, i);
            }
            __syncthreads();
            // ---------------------------------------------------- afterburner (:275-320)
            const int32_t ns = L.cnt[0], nb = L.cnt[1];
            auto burn = [&](vid_t i, int sub, int lanes) -> int64_t {
                const part_t p = P.part[i], best = P.dest[i];
                const int32_t ig = P.pregain[i];
                int64_t change = 0;
                const eoff_t re = g.rowptr[i + 1];
                for (eoff_t j = g.rowptr[i] + sub; j < re; j += lanes)
                    change += afterburn_edge(g.col[j], i, ig, g.ew[j], p, best, P.part, P.dest, P.pregain);
                return change;
            };
            {
                // small candidates: a group of P.lanes lanes each (the whole
                // workgroup takes the same number of rounds, so the group
                // sums and the list appends see every lane)
                const int G = P.lanes;
                const int gsub = t & (G - 1);
                const int32_t gid = t / G, ngroups = WG_BLOCK / G;
                for (int32_t x0 = 0; x0 < ns; x0 += ngroups) {
                    const int32_t x = x0 + gid;
                    const bool have = x < ns;
                    const vid_t i = have ? P.cand[x] : 0;
                    const int64_t change = group_sum(have ? burn(i, gsub, G) : 0, G);
                    const bool keep = have && gsub == 0 && (int64_t)P.pregain[i] + change >= 0;
                    if (keep) P.lock[i] = 1;
                    list_push(P.moves, n, &L.cnt[2], &L.cnt[3], keep, false, i);
                }
            }
            for (int32_t x = wv; x < nb; x += WG_WAVES) {
                const vid_t i = P.cand[n - 1 - x];
                const int64_t change = wave_reduce_sum(burn(i, lane, WAVE));
                if (lane == 0 && (int64_t)P.pregain[i] + change >= 0) {
                    P.lock[i] = 1;
                    P.moves[n - 1 - atomicAdd(&L.cnt[3], 1)] = i;
                }
            }
            __syncthreads();
        } else {
            // ---------------------------------------------------- rebalancing
            const bool strong = kind == 2;
            const int64_t max_dest = strong ? P.max_dest_strong : P.max_dest_weak;
            const int64_t smax = P.size_max;
            const int64_t width = (int64_t)MAX_BUCKETS * P.sections;
            const int32_t nbk = strong ? k * (int32_t)width + 1 : P.nbk;
            for (int32_t b = t; b < nbk; b += WG_BLOCK) sbk[b] = 0;
            if (wv == 0) {
                if (strong) {
                    // the largest receiver capacity (:440-448)
                    int64_t mx = INT64_MIN;
                    for (int32_t p = lane; p < k; p += WAVE) {
                        const int64_t sz = L.sizes[p];
                        if (sz < max_dest && max_dest - sz > mx) mx = max_dest - sz;
                    }
                    mx = wave_reduce_max(mx);
                    if (lane == 0) L.counter = mx;
                } else {
                    // undersized parts in increasing id (:578-591)
                    int32_t base = 0;
                    for (int32_t p0 = 0; p0 < k; p0 += WAVE) {
                        const int32_t p = p0 + lane;
                        const bool f = p < k && L.sizes[p] < max_dest;
                        const uint64_t mk = __ballot(f);
                        if (f) L.undersized[base + (int32_t)__popcll(mk & ((1ull << lane) - 1ull))] = p;
                        base += (int32_t)__popcll(mk);
                    }
                    if (lane == 0) L.nunder = base;
                }
            }
            for (int32_t p = t; p < k; p += WG_BLOCK) L.evict_adjust[p] = 0;
            __syncthreads();
            const int32_t tu = L.nunder;
            const int64_t max_vwgt = L.counter;
            // select (:592-643 weak, :449-481 strong)
            for (vid_t i = t; i < n; i += WG_BLOCK) {
                const part_t p = P.part[i];
                const int64_t sp = L.sizes[p];
                const int64_t w = g.vw[i];
                int32_t b = -1;
                part_t d = p;
                if (!strong) {
                    if (sp > smax && (double)w < 1.5 * (double)(sp - P.opt)) {
                        part_t best = p;
                        int32_t gain = 0, p_gain = 0;
                        for_slots(P.conn + P.conn_off[i], P.conn_size[i], [&](int2 e) {
                            if (e.x > NULL_PART && L.sizes[e.x] < max_dest) {
                                if (e.y > gain) {
                                    best = e.x;
                                    gain = e.y;
                                } else if (e.y == gain && gain > 0 && e.x < best) {
                                    best = e.x;
                                }
                            }
                            if (e.x == p) p_gain = e.y;
                        });
                        int64_t sv = 0;
                        if (gain > 0) {
                            d = best;
                            sv = (int64_t)gain - p_gain;
                        } else if (tu > 0) {
                            d = L.undersized[(uint32_t)i % (uint32_t)tu];
                            sv = -(int64_t)p_gain;
                        }
                        if (d != p) {
                            const int32_t gt = gain_bucket_fast(sv, w);
                            b = (MAX_BUCKETS * p + gt) * P.sections + (int32_t)((uint32_t)i % (uint32_t)P.sections);
                            atomicAdd(&sbk[b], (int32_t)w);
                        }
                    }
                    P.dest[i] = d;
                } else {
                    const bool vw_ok = max_vwgt > 0 && w <= 2 * max_vwgt;
                    if (sp > smax && vw_ok && w < 2 * (sp - P.opt)) {
                        uint64_t tk = 0, tg = 0;
                        int64_t p_gain = 0;
                        for_slots(P.conn + P.conn_off[i], P.conn_size[i], [&](int2 e) {
                            if (e.x == p) {
                                p_gain = e.y;
                            } else if (e.x > NULL_PART && L.sizes[e.x] < max_dest) {
                                tg += (uint64_t)e.y;
                                tk += 1;
                            }
                        });
                        if (tk == 0) tk = 1;
                        const int64_t gain = (int64_t)(tg / tk) - p_gain;
                        const int64_t vwgt = w < sp - smax ? w : sp - smax;
                        const int32_t gt = gain_bucket_fast(gain, vwgt);
                        if (gt < MAX_BUCKETS) {
                            b = (MAX_BUCKETS * p + gt) * P.sections + (int32_t)((uint32_t)i % (uint32_t)P.sections) + 1;
                            atomicAdd(&sbk[b], (int32_t)w);
                        }
                    }
                }
                P.bid[i] = b;
            }
            __syncthreads();
            block_scan_inplace(sbk, nbk, L.red);  // sbk = exclusive offsets (get_evictions, :363-384)
            // classification: whole buckets below a part's limit are evicted
            // (weak) or need their members' positions (strong); a weak bucket
            // straddling the limit needs them too
            // a weak bucket that straddles the limit of a part it holds
            // members of is ranked with all of its members: bucket 50 of
            // part q shares its index with part q + 1's first gain bucket,
            // and a member's position counts every member of lower id
            auto weak_straddle_of = [&](int32_t b, int32_t q) -> bool {
                const int64_t base = (int64_t)sbk[b] - sbk[(int64_t)q * width];
                const int64_t limit = L.sizes[q] - smax;
                return base < limit && base + ((int64_t)sbk[b + 1] - sbk[b]) > limit;
            };
            auto needs_rank = [&](vid_t i) -> bool {
                const int32_t b = P.bid[i];
                if (b < 0) return false;
                if (strong) {
                    const part_t p = P.part[i];
                    return (int64_t)sbk[b] - sbk[(int64_t)p * width] < L.sizes[p] - smax;
                }
                const int32_t q = (int32_t)(b / width);
                if (q < k && weak_straddle_of(b, q)) return true;
                return q >= 1 && (int64_t)b - (int64_t)q * width < P.sections && weak_straddle_of(b, q - 1);
            };
            if (!strong) {
                for (vid_t i = t; i < n; i += WG_BLOCK) {
                    const int32_t b = P.bid[i];
                    bool ev = false;
                    if (b >= 0) {
                        const part_t p = P.part[i];
                        const int64_t base = (int64_t)sbk[b] - sbk[(int64_t)p * width];
                        ev = base + ((int64_t)sbk[b + 1] - sbk[b]) <= L.sizes[p] - smax;
                    }
                    list_push(P.moves, n, &L.cnt[2], &L.cnt[3], ev, ev && deg_of(g, i) > P.big_deg, i);
                }
            }
            const int32_t R = ordered_compact_ids(n, P.rlist, needs_rank, L.red);
            if (R > WG_BLOCK)
                for (int32_t b = t; b < nbk; b += WG_BLOCK) P.runb[b] = 0;
            __syncthreads();
            // ranking by (bucket, id) in chunks of WG_BLOCK members in id order
            for (int32_t c0 = 0; c0 < R; c0 += WG_BLOCK) {
                const int32_t x = c0 + t;
                const bool valid = x < R;
                const vid_t v0 = valid ? P.rlist[x] : 0;
                uint64_t key = ~0ull;
                uint32_t w = 0;
                if (valid) {
                    key = ((uint64_t)(uint32_t)P.bid[v0] << 32) | (uint32_t)v0;
                    w = (uint32_t)g.vw[v0];
                }
                block_sort(key, w, L.key, L.pre);
                L.key[t] = key;
                int64_t tot;
                const int64_t ex = block_excl_scan<int64_t, WG_BLOCK>(key == ~0ull ? 0 : (int64_t)w, tot, L.red);
                L.pre[t] = ex;
                __syncthreads();
                bool ev = false, big = false;
                vid_t v = 0;
                if (key != ~0ull) {
                    const int32_t b = (int32_t)(key >> 32);
                    v = (vid_t)(uint32_t)key;
                    // first member of bucket b in this chunk
                    int32_t lo = 0, hi = t;
                    const uint64_t kb = (uint64_t)(uint32_t)b << 32;
                    while (lo < hi) {
                        const int32_t mid = (lo + hi) >> 1;
                        if (L.key[mid] < kb) lo = mid + 1;
                        else hi = mid;
                    }
                    const int64_t prior = c0 > 0 ? (int64_t)P.runb[b] : 0;
                    const int64_t pos = prior + ex - L.pre[lo];
                    const part_t p = P.part[v];
                    const int64_t base = (int64_t)sbk[b] - sbk[(int64_t)p * width];
                    const int64_t limit = L.sizes[p] - smax;
                    // a weak member decides here only when its own part's
                    // limit straddles the bucket (whole buckets were decided
                    // by the classification)
                    const bool own = strong || weak_straddle_of(b, p);
                    if (own && base + pos < limit) {
                        ev = true;
                        big = deg_of(g, v) > P.big_deg;
                        if (strong) {
                            if (base + pos + (int64_t)w >= limit) L.evict_adjust[p] = base + pos + (int64_t)w;
                            P.score[v] = (int64_t)sbk[b] + pos;
                        }
                    } else if (strong) {
                        P.dest[v] = p;  // kept: not a re-draw candidate below
                    }
                    // the chunk's last member of bucket b carries its total forward
                    const bool last = t + 1 == WG_BLOCK || L.key[t + 1] == ~0ull || (int32_t)(L.key[t + 1] >> 32) != b;
                    if (last && R > WG_BLOCK) P.runb[b] = (int32_t)(prior + ex + (int64_t)w - L.pre[lo]);
                }
                list_push(P.moves, n, &L.cnt[2], &L.cnt[3], ev, big, v);
                __syncthreads();
            }
            if (strong) {
                // cookie cutter (:491-518): two scans over the parts, by wave 0
                if (wv == 0) {
                    int64_t run = 0, run2 = 0;
                    for (int32_t p0 = 0; p0 < k; p0 += WAVE) {
                        const int32_t p = p0 + lane;
                        int64_t add = 0, cap = 0;
                        if (p < k) {
                            const int64_t begin = (int64_t)MAX_BUCKETS * p * P.sections;
                            add = L.evict_adjust[p];
                            if (add == 0) add = (int64_t)sbk[begin + width] - sbk[begin];
                            cap = max_dest > L.sizes[p] ? max_dest - L.sizes[p] : 0;
                        }
                        const int64_t inc = wave_incl_scan(add), inc2 = wave_incl_scan(cap);
                        if (p < k) {
                            const int64_t begin = (int64_t)MAX_BUCKETS * p * P.sections;
                            L.evict_adjust[p] = (int64_t)sbk[begin] - (run + inc - add);
                            L.evict_start[p + 1] = run2 + inc2;
                        }
                        run += __shfl(inc, WAVE - 1, WAVE);
                        run2 += __shfl(inc2, WAVE - 1, WAVE);
                    }
                    if (lane == 0) {
                        L.evict_start[0] = 0;
                        L.counter = run;
                    }
                }
                __syncthreads();
                // destinations (:519-544): a member takes the chunk that holds at
                // least half of its weight; the others are re-drawn below
                const int32_t ns = L.cnt[2], nb = L.cnt[3];
                for (int32_t x = t; x < ns + nb; x += WG_BLOCK) {
                    const vid_t v = x < ns ? P.moves[x] : P.moves[n - 1 - (x - ns)];
                    const part_t pv = P.part[v];
                    const int64_t s = P.score[v] - L.evict_adjust[pv];
                    P.score[v] = s;
                    const int32_t c = walk_chunks(L.evict_start, 0, k + 1, s) - 1;
                    const int64_t w = g.vw[v];
                    part_t d = pv;
                    if (c < k && w / 2 <= L.evict_start[c + 1] - s) d = c;
                    else if (c < k) d = NO_MOVE;  // re-drawn in id order
                    P.dest[v] = d;
                }
                __syncthreads();
                // the re-drawn members in increasing id (the ranked list's order)
                const int32_t nr =
                    ordered_compact(P.rlist, R, P.cand, [&](vid_t v) { return P.dest[v] == NO_MOVE; }, L.red);
                if (t == 0) {
                    int64_t counter = L.counter;
                    for (int32_t r = 0; r < nr; r++) {
                        const vid_t v = P.cand[r];
                        const int64_t w = g.vw[v];
                        int64_t s = P.score[v];
                        int32_t p = 0;
                        part_t d = P.part[v];
                        // every pass with p < k re-draws (s grows) or stops,
                        // so at most k + 1 passes reach a chunk past the last
                        for (int32_t guard = 0; p < k; guard++) {
                            if (guard > 2 * k + 2) {
                                atomicOr(err, 8);
                                break;
                            }
                            p = walk_chunks(L.evict_start, p, k + 1, s) - 1;
                            if (p < k && w / 2 <= L.evict_start[p + 1] - s) {
                                d = p;
                                break;
                            } else if (p < k) {
                                s = counter;
                                counter += w;
                            }
                        }
                        P.dest[v] = d;
                    }
                }
                __syncthreads();
            }
        }

        // -------------------------------------------------------- perform_moves (:905-955)
        const int32_t ms = L.cnt[2], mb = L.cnt[3];
        const bool rebuild = ms + mb > n / 10;  // update_large (:933)
        // the cut change before the moves: for each edge (v, u) of a move,
        // w * ([part(u) == new] - [part(u) == old]); part sizes; cached
        // destinations of the moved vertices cleared
        auto cut_side = [&](vid_t v, part_t pnew, part_t pold, int sub, int lanes) -> int64_t {
            int64_t c = 0;
            const eoff_t re = g.rowptr[v + 1];
            for (eoff_t j = g.rowptr[v] + sub; j < re; j += lanes) {
                const part_t pu = P.part[g.col[j]];
                c += (int64_t)g.ew[j] * ((pu == pnew ? 1 : 0) - (pu == pold ? 1 : 0));
            }
            return c;
        };
        int64_t cc = over_items(P.moves, n, ms, mb, P.lanes, [&](vid_t v, int sub, int lanes) -> int64_t {
            const part_t pold = P.part[v], pnew = P.dest[v];
            if (sub == 0) {
                P.dest_cache[v] = NULL_PART;
                const unsigned long long w = (unsigned long long)g.vw[v];
                atomicAdd((unsigned long long*)&L.sizes[pold], (unsigned long long)(-(long long)w));
                atomicAdd((unsigned long long*)&L.sizes[pnew], w);
            }
            return cut_side(v, pnew, pold, sub, lanes);
        });
        __syncthreads();
        // the moves; dest keeps the old part (:926-929)
        for (int32_t x = t; x < ms + mb; x += WG_BLOCK) {
            const vid_t v = x < ms ? P.moves[x] : P.moves[n - 1 - (x - ms)];
            const part_t pold = P.part[v];
            P.part[v] = P.dest[v];
            P.dest[v] = pold;
        }
        __syncthreads();
        // the cut change after the moves; update_small's subtractions
        // (:785-812) or the rows to rebuild marked (update_large, :743-756)
        cc += over_items(P.moves, n, ms, mb, P.lanes, [&](vid_t v, int sub, int lanes) -> int64_t {
            const part_t pnew = P.part[v], pold = P.dest[v];
            int64_t c = 0;
            const eoff_t re = g.rowptr[v + 1];
            for (eoff_t j = g.rowptr[v] + sub; j < re; j += lanes) {
                const vid_t u = g.col[j];
                const int32_t w = g.ew[j];
                const part_t pu = P.part[u];
                c += (int64_t)w * ((pu == pnew ? 1 : 0) - (pu == pold ? 1 : 0));
                if (rebuild) P.mark[u] = 1;
                else conn_sub(P.conn + P.conn_off[u], P.conn_size[u], pold, w, k, err);
            }
            return c;
        });
        __syncthreads();
        if (!rebuild) {
            // update_small's additions (:815-886)
            over_items(P.moves, n, ms, mb, P.lanes, [&](vid_t v, int sub, int lanes) -> int64_t {
                const part_t pnew = P.part[v];
                const eoff_t re = g.rowptr[v + 1];
                for (eoff_t j = g.rowptr[v] + sub; j < re; j += lanes) {
                    const vid_t u = g.col[j];
                    P.dest_cache[u] = NULL_PART;
                    const eoff_t co = P.conn_off[u];
                    conn_add(P.conn + co, &P.conn_size[u], (int32_t)(P.conn_off[u + 1] - co), pnew, g.ew[j], err);
                }
                return 0;
            });
        }
        cc = block_reduce_sum<int64_t, WG_BLOCK>(cc, L.red);
        int64_t mx = INT64_MIN;
        for (int32_t p = t; p < k; p += WG_BLOCK) mx = L.sizes[p] > mx ? L.sizes[p] : mx;
        mx = block_reduce_max<int64_t, WG_BLOCK>(mx, L.red);

        // -------------------------------------------------------- best state (:1117-1132)
        if (t == 0) {
            WgState& S = L.S;
            S.curr_cut -= cc;
            S.curr_imb = mx > P.opt ? mx - P.opt : 0;
            if (kind == 0) {
                S.balance_counter = 0;
                S.lp_iterations++;
            } else {
                S.balance_counter++;
            }
            int take = 0;
            if (S.best_imb > P.imb_max && S.curr_imb < S.best_imb) {
                take = 1;
                S.count = 0;
            } else if (S.curr_cut < S.best_cut && (S.curr_imb <= P.imb_max || S.curr_imb <= S.best_imb)) {
                if ((double)S.curr_cut < P.tol * (double)S.best_cut) S.count = 0;
                take = 2;
            }
            if (take) {
                S.best_cut = S.curr_cut;
                S.best_imb = S.curr_imb;
                S.changed = 1;
            }
            L.take = take;
            L.rebuild = rebuild ? 1 : 0;
        }
        __syncthreads();
        if (L.take) {
            for (vid_t i = t; i < n; i += WG_BLOCK) P.best_part[i] = P.part[i];
            for (int32_t p = t; p < k; p += WG_BLOCK) P.best_sizes[p] = L.sizes[p];
        }
        if (L.rebuild) {
            if (t == 0) L.S.status = WG_REBUILD;
            __syncthreads();
            break;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int32_t p = t; p < k; p += WG_BLOCK) P.curr_sizes[p] = L.sizes[p];
    if (t == 0) {
        L.S.err = ld_rx(err);  // table errors the steps reported (atomics, not cached)
        *P.st = L.S;
        if (mirror) *mirror = L.S;
    }
}

__global__ void __launch_bounds__(WG_BLOCK) k_refine_wg(WgParams P) { refine_wg_body(P, nullptr); }

// Graph-tagged: workgroup b runs the level that ps[b] describes (levels of
// different graphs, each with its own arrays), and copies its final state to
// states[b].
__global__ void __launch_bounds__(WG_BLOCK) k_refine_wg_multi(const WgParams* ps, WgState* states) {
    refine_wg_body(ps[blockIdx.x], states + blockIdx.x);
}

__global__ void k_wg_init(WgState* st, int64_t cut, int64_t imb) {
    WgState S{};
    S.curr_cut = S.best_cut = cut;
    S.curr_imb = S.best_imb = imb;
    *st = S;
}

void wg_init_state(WgState* st, int64_t cut, int64_t imb, hipStream_t s) {
    k_wg_init<<<1, 1, 0, s>>>(st, cut, imb);
    JET_LAUNCH_CHECK();
}

// The dynamic LDS of the bucket array may take the kernels past 64 KB
// (gfx950 has 160 KB per workgroup): allowed once per device and kernel.
static void allow_wg_lds(const void* kernel, int which) {
    static std::once_flag once[2][64];
    static hipError_t attr_rc[2][64];
    int dev = 0;
    JET_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw DeviceError("device index beyond 64");
    std::call_once(once[which][dev], [&] {
        attr_rc[which][dev] = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)(sizeof(int32_t) * (WG_MAX_NBK + 1)));
    });
    JET_HIP(attr_rc[which][dev]);
}

static void check_wg_params(const WgParams& p) {
    if (p.nbk < 1 || p.nbk > WG_MAX_NBK || p.k < 1 || p.k > WG_MAX_K || p.ntemps > WG_MAX_TEMPS)
        throw DeviceError("one-workgroup refinement: level outside the kernel's limits");
}

void launch_refine_wg(const WgParams& p, hipStream_t s) {
    allow_wg_lds(reinterpret_cast<const void*>(&k_refine_wg), 0);
    check_wg_params(p);
    const size_t dyn = sizeof(int32_t) * (size_t)(p.nbk + 1);
    k_refine_wg<<<1, WG_BLOCK, dyn, s>>>(p);
    JET_LAUNCH_CHECK();
}

void launch_refine_wg_multi(const WgParams* host_ps, const WgParams* dev_ps, WgState* dev_states, int count,
                            hipStream_t s) {
    if (count <= 0) return;
    allow_wg_lds(reinterpret_cast<const void*>(&k_refine_wg_multi), 1);
    int32_t nbk = 1;
    for (int i = 0; i < count; i++) {
        check_wg_params(host_ps[i]);
        nbk = std::max(nbk, host_ps[i].nbk);
    }
    const size_t dyn = sizeof(int32_t) * (size_t)(nbk + 1);
    k_refine_wg_multi<<<count, WG_BLOCK, dyn, s>>>(dev_ps, dev_states);
    JET_LAUNCH_CHECK();
}

}  // namespace jet
