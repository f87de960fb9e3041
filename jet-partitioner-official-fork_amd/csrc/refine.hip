// Jet refinement for gfx950 (reference: src/jet_refiner.hpp, src/part_stat.hpp,
// src/uncoarsen.hpp).
//
// Connectivity tables keep the reference's mechanics (capacity min(deg, k),
// direct mapping when the table size is k, linear probing from p % size
// otherwise, shrink to used + max(used/4, 3), HASH_RECLAIM slots, dense
// overflow append) so their contents, sizes and zero entries evolve exactly
// as in the reference. Slot order inside a table is an accident of insertion
// timing on every parallel backend; every decision that reads a table here is
// made independent of it: among parts with equal connectivity the lowest part
// id wins (for direct-mapped tables that is the reference's first-in-slot
// rule). Rebalancer minibucket positions follow increasing vertex id (a stable
// sort by bucket id), and the strong rebalancer's destination re-draws run in
// increasing vertex id.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "prims.h"
#include "refine.h"
#include "wave.h"

namespace jet {

constexpr int32_t SMALL_DEG = 16;       // rows built by one lane
constexpr int32_t WAVE_TABLE_MAX = 1024; // rows with capacity <= this built by one wave
constexpr int32_t BIG_DEG = 64;         // candidates/moves above this degree use one wave
constexpr int32_t MAX_K = 8192;
constexpr int32_t LDS_K = 2048;         // part-size deltas aggregated in LDS up to this k
constexpr int32_t HEAVY_DEG_REFINE = 2048;  // rows above this degree are built by a workgroup

// ---------------------------------------------------------------- helpers
__device__ inline int32_t ld_rx(const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline int32_t lookup(const int2* tab, int32_t size, part_t target) {
    for (int32_t q = 0; q < size; q++) {
        int32_t pi = (target + q) % size;
        int2 e = tab[pi];
        if (e.x == target) return e.y;
        if (e.x == NULL_PART) return 0;
    }
    return 0;
}

__device__ inline void push_list(vid_t* list, int32_t* counters, int idx_small, vid_t nmax, vid_t v, eoff_t deg) {
    if (deg > BIG_DEG) list[nmax - 1 - atomicAdd(&counters[idx_small + 1], 1)] = v;
    else list[atomicAdd(&counters[idx_small], 1)] = v;
}

__device__ inline vid_t list_at(const vid_t* list, int32_t nsmall, vid_t nmax, int64_t x) {
    return x < nsmall ? list[x] : list[nmax - 1 - (x - nsmall)];
}

// ---------------------------------------------------------------- statistics
__global__ void k_cut2(DGraph g, const part_t* part, unsigned long long* out) {
    __shared__ int64_t red[256 / WAVE];
    int64_t c = 0;
    FOR_THREAD(i, g.n) {
        part_t p = part[i];
        for (eoff_t j = g.rowptr[i]; j < g.rowptr[i + 1]; j++)
            if (part[g.col[j]] != p) c += g.ew[j];
    }
    c = block_reduce_sum<int64_t, 256>(c, red);
    if (threadIdx.x == 0 && c) atomicAdd(out, (unsigned long long)c);
}

int64_t total_cut2(const DGraph& g, const part_t* part, Arena& a, hipStream_t s) {
    ArenaScope sc(a);
    unsigned long long* out = a.alloc<unsigned long long>(1);
    fill_i64(reinterpret_cast<int64_t*>(out), 0, 1, s);
    if (g.n > 0) {
        k_cut2<<<grid_for(g.n, 256, 8192), 256, 0, s>>>(g, part, out);
        JET_LAUNCH_CHECK();
    }
    return (int64_t)read_scalar(out, s);
}

__global__ void k_part_sizes(DGraph g, const part_t* part, int32_t k, unsigned long long* sizes) {
    __shared__ unsigned long long ls[LDS_K];
    bool lds = k <= LDS_K;
    if (lds) {
        for (int t = threadIdx.x; t < k; t += blockDim.x) ls[t] = 0;
        __syncthreads();
    }
    FOR_THREAD(i, g.n) {
        if (lds) atomicAdd(&ls[part[i]], (unsigned long long)g.vw[i]);
        else atomicAdd(&sizes[part[i]], (unsigned long long)g.vw[i]);
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < k; t += blockDim.x)
            if (ls[t]) atomicAdd(&sizes[t], ls[t]);
    }
}

void part_sizes(const DGraph& g, const part_t* part, int32_t k, int64_t* sizes, hipStream_t s) {
    fill_i64(sizes, 0, k, s);
    if (g.n > 0) {
        k_part_sizes<<<grid_for(g.n, 256, 2048), 256, 0, s>>>(g, part, k, reinterpret_cast<unsigned long long*>(sizes));
        JET_LAUNCH_CHECK();
    }
}

__global__ void k_project(vid_t n, const vid_t* map, const part_t* coarse, part_t* fine) {
    FOR_THREAD(i, n) fine[i] = coarse[map[i]];
}

void project(vid_t n_fine, const vid_t* map, const part_t* coarse, part_t* fine, hipStream_t s) {
    if (n_fine <= 0) return;
    k_project<<<grid_for(n_fine, 256, 16384), 256, 0, s>>>(n_fine, map, coarse, fine);
    JET_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- table build
// jet_refiner.hpp:958-1048 and :648-739: capacity min(deg, k); direct slot p
// when the capacity is k, else linear probing from p % cap; shrink to
// used + max(used/4, 3) when that is smaller. The whole capacity is written
// (layout followed by empty slots). `mark` (optional) restricts the build to
// marked rows, clears their mark and their destination cache (update_large,
// :743-777).

__global__ void k_conn_caps(DGraph g, int32_t k, int32_t* caps) {
    FOR_THREAD(i, g.n) {
        int64_t d = g.rowptr[i + 1] - g.rowptr[i];
        caps[i] = (int32_t)(d > k ? k : d);
    }
}

__global__ void k_conn_tiers(DGraph g, int32_t k, vid_t* mid, vid_t* big, int32_t* cnt) {
    FOR_THREAD(i, g.n) {
        int64_t d = g.rowptr[i + 1] - g.rowptr[i];
        if (d <= SMALL_DEG) continue;
        int64_t cap = d > k ? k : d;
        if (d <= HEAVY_DEG_REFINE && cap <= WAVE_TABLE_MAX) mid[atomicAdd(&cnt[0], 1)] = (vid_t)i;
        else big[atomicAdd(&cnt[1], 1)] = (vid_t)i;
    }
}

__device__ inline int32_t shrink_size(int32_t used, int32_t cap) {
    int32_t q = used / 4;
    if (q < 3) q = 3;
    int32_t sz = used + q;
    return sz < cap ? sz : cap;
}

constexpr int SB_BLOCK = 256;

__global__ void __launch_bounds__(SB_BLOCK)
    k_conn_build_small(DGraph g, const part_t* part, int32_t k, const eoff_t* conn_off, int2* conn,
                       int32_t* conn_size, int32_t* mark, part_t* dest_cache) {
    __shared__ int2 tab[SMALL_DEG][SB_BLOCK];
    const int t = threadIdx.x;
    FOR_THREAD(i, g.n) {
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        int32_t deg = (int32_t)(re - rs);
        if (deg > SMALL_DEG) continue;
        if (mark) {
            if (!mark[i]) continue;
            mark[i] = 0;
            dest_cache[i] = NULL_PART;
        }
        int32_t cap = deg < k ? deg : k;
        for (int32_t s = 0; s < cap; s++) tab[s][t] = make_int2(NULL_PART, 0);
        int32_t used = 0;
        for (eoff_t j = rs; j < re; j++) {
            part_t p = part[g.col[j]];
            int32_t po = p % cap;
            if (cap < k) {
                while (tab[po][t].x != NULL_PART && tab[po][t].x != p) po = (po + 1) % cap;
            }
            if (tab[po][t].x == NULL_PART) {
                tab[po][t].x = p;
                used++;
            }
            tab[po][t].y += g.ew[j];
        }
        int2* out = conn + conn_off[i];
        int32_t size = shrink_size(used, cap);
        if (size < cap) {
            for (int32_t s = 0; s < cap; s++) out[s] = make_int2(NULL_PART, 0);
            for (int32_t s = 0; s < cap; s++) {
                int2 e = tab[s][t];
                if (e.x == NULL_PART) continue;
                int32_t po = e.x % size;
                while (out[po].x != NULL_PART) po = (po + 1) % size;
                out[po] = e;
            }
        } else {
            for (int32_t s = 0; s < cap; s++) out[s] = tab[s][t];
        }
        conn_size[i] = size;
    }
}

// Inserts part p with weight w into an LDS table of `cap` slots (cap == k:
// direct; else probing); counts newly used slots.
__device__ inline void lds_table_add(int2* T, int32_t cap, int32_t k, part_t p, int32_t w, int32_t* used) {
    int32_t po = p % cap;
    if (cap == k) {
        if (T[po].x == NULL_PART && atomicCAS(&T[po].x, NULL_PART, p) == NULL_PART) atomicAdd(used, 1);
    } else {
        for (int32_t guard = 0; guard <= cap; guard++) {
            int32_t px = T[po].x;
            if (px == p) break;
            if (px == NULL_PART) {
                int32_t old = atomicCAS(&T[po].x, NULL_PART, p);
                if (old == NULL_PART) {
                    atomicAdd(used, 1);
                    break;
                }
                if (old == p) break;
            }
            po = (po + 1) % cap;
        }
    }
    atomicAdd(&T[po].y, w);
}

// Re-inserts the entries of T (cap slots) into T2 (size slots, linear probing).
__device__ inline void lds_table_shrink(const int2* T, int32_t cap, int2* T2, int32_t size, int start, int stride) {
    for (int32_t s = start; s < cap; s += stride) {
        int2 e = T[s];
        if (e.x == NULL_PART) continue;
        int32_t po = e.x % size;
        for (int32_t guard = 0; guard <= size; guard++) {
            if (T2[po].x == NULL_PART && atomicCAS(&T2[po].x, NULL_PART, e.x) == NULL_PART) {
                T2[po].y = e.y;
                break;
            }
            po = (po + 1) % size;
        }
    }
}

constexpr int WB_WAVES = 4;

__global__ void __launch_bounds__(WB_WAVES* WAVE)
    k_conn_build_wave(DGraph g, const part_t* part, int32_t k, const eoff_t* conn_off, int2* conn,
                      int32_t* conn_size, int32_t* mark, part_t* dest_cache, const vid_t* list, int32_t nlist) {
    __shared__ int2 T[WB_WAVES][WAVE_TABLE_MAX];
    __shared__ int2 T2[WB_WAVES][WAVE_TABLE_MAX];
    __shared__ int32_t used[WB_WAVES];
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    for (int32_t r = blockIdx.x * WB_WAVES + w; r < nlist; r += gridDim.x * WB_WAVES) {
        vid_t i = list[r];
        if (mark) {
            if (!mark[i]) continue;
        }
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        int32_t deg = (int32_t)(re - rs);
        int32_t cap = deg < k ? deg : k;
        for (int32_t s = lane; s < cap; s += WAVE) {
            T[w][s] = make_int2(NULL_PART, 0);
            T2[w][s] = make_int2(NULL_PART, 0);
        }
        if (lane == 0) used[w] = 0;
        wave_sync();
        for (eoff_t j = rs + lane; j < re; j += WAVE) lds_table_add(T[w], cap, k, part[g.col[j]], g.ew[j], &used[w]);
        wave_sync();
        int32_t size = shrink_size(used[w], cap);
        int2* out = conn + conn_off[i];
        if (size < cap) {
            lds_table_shrink(T[w], cap, T2[w], size, lane, WAVE);
            wave_sync();
            for (int32_t s = lane; s < cap; s += WAVE) out[s] = T2[w][s];
        } else {
            for (int32_t s = lane; s < cap; s += WAVE) out[s] = T[w][s];
        }
        if (lane == 0) {
            conn_size[i] = size;
            if (mark) {
                mark[i] = 0;
                dest_cache[i] = NULL_PART;
            }
        }
        wave_sync();
    }
}

constexpr int BB_BLOCK = 512;

__global__ void __launch_bounds__(BB_BLOCK)
    k_conn_build_block(DGraph g, const part_t* part, int32_t k, const eoff_t* conn_off, int2* conn,
                       int32_t* conn_size, int32_t* mark, part_t* dest_cache, const vid_t* list, int32_t nlist) {
    __shared__ int2 T[MAX_K];
    __shared__ int2 T2[MAX_K];
    __shared__ int32_t used;
    for (int32_t r = blockIdx.x; r < nlist; r += gridDim.x) {
        vid_t i = list[r];
        if (mark && !mark[i]) continue;
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        int64_t deg = re - rs;
        int32_t cap = (int32_t)(deg < k ? deg : k);
        for (int32_t s = threadIdx.x; s < cap; s += BB_BLOCK) {
            T[s] = make_int2(NULL_PART, 0);
            T2[s] = make_int2(NULL_PART, 0);
        }
        if (threadIdx.x == 0) used = 0;
        __syncthreads();
        for (eoff_t j = rs + threadIdx.x; j < re; j += BB_BLOCK) lds_table_add(T, cap, k, part[g.col[j]], g.ew[j], &used);
        __syncthreads();
        int32_t size = shrink_size(used, cap);
        int2* out = conn + conn_off[i];
        if (size < cap) {
            lds_table_shrink(T, cap, T2, size, threadIdx.x, BB_BLOCK);
            __syncthreads();
            for (int32_t s = threadIdx.x; s < cap; s += BB_BLOCK) out[s] = T2[s];
        } else {
            for (int32_t s = threadIdx.x; s < cap; s += BB_BLOCK) out[s] = T[s];
        }
        if (threadIdx.x == 0) {
            conn_size[i] = size;
            if (mark) {
                mark[i] = 0;
                dest_cache[i] = NULL_PART;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- jet_lp
// jet_refiner.hpp:211-268: destination selection, viable-move filter.
__global__ void k_lp_select(DGraph g, const part_t* part, const eoff_t* conn_off, const int2* conn,
                            const int32_t* conn_size, part_t* dest_cache, int32_t* gain_p, int32_t* lock, double fr,
                            part_t* dest, int32_t* pregain, vid_t* cand, int32_t* counters, vid_t nmax) {
    FOR_THREAD(i, g.n) {
        part_t best = dest_cache[i];
        if (best == NULL_PART) {
            part_t p = part[i];
            best = NO_MOVE;
            int32_t b_conn = 0, p_conn = 0;
            const int2* tab = conn + conn_off[i];
            int32_t sz = conn_size[i];
            for (int32_t s = 0; s < sz; s++) {
                int2 e = tab[s];
                if (e.y > b_conn && e.x != p) {
                    best = e.x;
                    b_conn = e.y;
                } else if (e.y == b_conn && b_conn > 0 && e.x != p && e.x < best) {
                    best = e.x;
                } else if (e.y > 0 && e.x == p) {
                    p_conn = e.y;
                }
            }
            int32_t gsv = 0;
            if (best != NO_MOVE) {
                if (b_conn >= p_conn || ((double)(p_conn - b_conn) < floor(fr * (double)p_conn))) gsv = b_conn - p_conn;
                else best = NO_MOVE;
            }
            gain_p[i] = gsv;
            dest_cache[i] = best;
        }
        dest[i] = best;
        if (best != NO_MOVE && lock[i] == 0) {
            pregain[i] = gain_p[i];
            push_list(cand, counters, 0, nmax, (vid_t)i, g.rowptr[i + 1] - g.rowptr[i]);
        } else {
            pregain[i] = GAIN_MIN;
            lock[i] = 0;
        }
    }
}

// jet_refiner.hpp:275-307 (afterburner): gain correction assuming every
// higher-priority neighbour (larger pre-gain, or equal and smaller id) moves.
__device__ inline int64_t afterburn_edge(vid_t v, vid_t i, int32_t ig, int32_t w, part_t p, part_t best,
                                         const part_t* part, const part_t* dest, const int32_t* pregain) {
    int32_t vg = pregain[v];
    int64_t ch = 0;
    if (vg > ig || (vg == ig && v < i)) {
        part_t vp = dest[v];
        if (vp == p) ch -= w;
        else if (vp == best) ch += w;
        vp = part[v];
        if (vp == p) ch += w;
        else if (vp == best) ch -= w;
    }
    return ch;
}

__global__ void k_afterburner_small(DGraph g, const vid_t* cand, const int32_t* counters, const part_t* part,
                                    const part_t* dest, const int32_t* pregain, int32_t* lock, vid_t* moves,
                                    int32_t* mcounters, vid_t nmax) {
    const int32_t nc = counters[0];
    FOR_THREAD(x, nc) {
        vid_t i = cand[x];
        part_t p = part[i], best = dest[i];
        int32_t ig = pregain[i];
        int64_t change = 0;
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        for (eoff_t j = rs; j < re; j++) change += afterburn_edge(g.col[j], i, ig, g.ew[j], p, best, part, dest, pregain);
        if ((int64_t)ig + change >= 0) {
            lock[i] = 1;
            push_list(moves, mcounters, 2, nmax, i, re - rs);
        }
    }
}

__global__ void k_afterburner_big(DGraph g, const vid_t* cand, const int32_t* counters, const part_t* part,
                                  const part_t* dest, const int32_t* pregain, int32_t* lock, vid_t* moves,
                                  int32_t* mcounters, vid_t nmax) {
    const int32_t nb = counters[1];
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t wstride = ((int64_t)gridDim.x * blockDim.x) / WAVE;
    for (int64_t x = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; x < nb; x += wstride) {
        vid_t i = cand[nmax - 1 - x];
        part_t p = part[i], best = dest[i];
        int32_t ig = pregain[i];
        int64_t change = 0;
        eoff_t rs = g.rowptr[i], re = g.rowptr[i + 1];
        for (eoff_t j = rs + lane; j < re; j += WAVE)
            change += afterburn_edge(g.col[j], i, ig, g.ew[j], p, best, part, dest, pregain);
        change = wave_reduce_sum(change);
        if (lane == 0 && (int64_t)ig + change >= 0) {
            lock[i] = 1;
            push_list(moves, mcounters, 2, nmax, i, re - rs);
        }
    }
}

// ---------------------------------------------------------------- perform_moves
// jet_refiner.hpp:905-955.
__global__ void k_moves_apply(const vid_t* moves, const int32_t* mc, vid_t nmax, const wgt_t* vw, part_t* part,
                              part_t* dest, part_t* dest_cache, const eoff_t* conn_off, const int2* conn,
                              const int32_t* conn_size, unsigned long long* sizes, int32_t k, int64_t* acc) {
    __shared__ int64_t red[256 / WAVE];
    __shared__ unsigned long long ls[LDS_K];
    const bool lds = k <= LDS_K;
    if (lds) {
        for (int t = threadIdx.x; t < k; t += blockDim.x) ls[t] = 0;
        __syncthreads();
    }
    const int32_t ns = mc[2], total = mc[2] + mc[3];
    int64_t cc = 0;
    FOR_THREAD(x, total) {
        vid_t v = list_at(moves, ns, nmax, x);
        part_t p = part[v], b = dest[v];
        const int2* tab = conn + conn_off[v];
        int32_t sz = conn_size[v];
        cc += (int64_t)lookup(tab, sz, b) - lookup(tab, sz, p);
        dest_cache[v] = NULL_PART;
        unsigned long long w = (unsigned long long)vw[v];
        if (lds) {
            atomicAdd(&ls[p], (unsigned long long)(-(long long)w));
            atomicAdd(&ls[b], w);
        } else {
            atomicAdd(&sizes[p], (unsigned long long)(-(long long)w));
            atomicAdd(&sizes[b], w);
        }
        part[v] = b;
        dest[v] = p;
    }
    cc = block_reduce_sum<int64_t, 256>(cc, red);
    if (threadIdx.x == 0 && cc) atomicAdd((unsigned long long*)&acc[0], (unsigned long long)cc);
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < k; t += blockDim.x)
            if (ls[t]) atomicAdd(&sizes[t], ls[t]);
    }
}

// update_small subtract (jet_refiner.hpp:785-812).
__device__ inline void conn_sub(int2* tab, int32_t vsz, part_t p, int32_t w, int32_t k, int32_t* err) {
    int32_t po = p % vsz;
    int32_t guard = 0;
    while (ld_rx(&tab[po].x) != p) {
        po = (po + 1) % vsz;
        if (++guard > vsz) {
            atomicOr(err, 1);
            return;
        }
    }
    int32_t x = atomicAdd(&tab[po].y, -w);
    if (vsz < k && x == w) __hip_atomic_store(&tab[po].x, HASH_RECLAIM, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// update_small add (jet_refiner.hpp:815-886).
__device__ inline void conn_add(int2* tab, int32_t* vsize, int32_t cap, part_t best, int32_t w, int32_t* err) {
    int32_t vsz = ld_rx(vsize);
    int32_t po = best % vsz;
    bool success = false;
    for (int32_t q = 0; q < vsz; q++) {
        int32_t pi = (best + q) % vsz;
        int32_t e = ld_rx(&tab[pi].x);
        if (e == best) {
            success = true;
            po = pi;
            break;
        } else if (e == NULL_PART) {
            break;
        }
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

template <bool ADD>
__device__ inline void update_edge(vid_t v, int32_t w, part_t q, const eoff_t* conn_off, int2* conn,
                                   int32_t* conn_size, part_t* dest_cache, int32_t k, int32_t* err) {
    int2* tab = conn + conn_off[v];
    if (ADD) {
        dest_cache[v] = NULL_PART;
        conn_add(tab, &conn_size[v], (int32_t)(conn_off[v + 1] - conn_off[v]), q, w, err);
    } else {
        conn_sub(tab, ld_rx(&conn_size[v]), q, w, k, err);
    }
}

// ADD == false: subtract from the old part (held in dest); ADD == true: add to
// the new part (held in part).
template <bool ADD>
__global__ void k_update_small(DGraph g, const vid_t* moves, const int32_t* mc, vid_t nmax, const part_t* part,
                               const part_t* dest, const eoff_t* conn_off, int2* conn, int32_t* conn_size,
                               part_t* dest_cache, int32_t k, int32_t* err) {
    const int32_t ns = mc[2];
    FOR_THREAD(x, ns) {
        vid_t i = moves[x];
        part_t q = ADD ? part[i] : dest[i];
        for (eoff_t j = g.rowptr[i]; j < g.rowptr[i + 1]; j++)
            update_edge<ADD>(g.col[j], g.ew[j], q, conn_off, conn, conn_size, dest_cache, k, err);
    }
}

template <bool ADD>
__global__ void k_update_big(DGraph g, const vid_t* moves, const int32_t* mc, vid_t nmax, const part_t* part,
                             const part_t* dest, const eoff_t* conn_off, int2* conn, int32_t* conn_size,
                             part_t* dest_cache, int32_t k, int32_t* err) {
    const int32_t nb = mc[3];
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t wstride = ((int64_t)gridDim.x * blockDim.x) / WAVE;
    for (int64_t x = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; x < nb; x += wstride) {
        vid_t i = moves[nmax - 1 - x];
        part_t q = ADD ? part[i] : dest[i];
        for (eoff_t j = g.rowptr[i] + lane; j < g.rowptr[i + 1]; j += WAVE)
            update_edge<ADD>(g.col[j], g.ew[j], q, conn_off, conn, conn_size, dest_cache, k, err);
    }
}

// update_large mark step (jet_refiner.hpp:748-755).
__global__ void k_mark_adjacent(DGraph g, const vid_t* moves, const int32_t* mc, vid_t nmax, int32_t* mark) {
    const int32_t total = mc[2] + mc[3];
    FOR_THREAD(x, total) {
        vid_t i = list_at(moves, mc[2], nmax, x);
        for (eoff_t j = g.rowptr[i]; j < g.rowptr[i + 1]; j++) mark[g.col[j]] = 1;
    }
}

// cut change part 2 and the largest part (jet_refiner.hpp:937-948).
__global__ void k_moves_finish(const vid_t* moves, const int32_t* mc, vid_t nmax, const part_t* part,
                               const part_t* dest, const eoff_t* conn_off, const int2* conn, const int32_t* conn_size,
                               const int64_t* sizes, int32_t k, int64_t* acc) {
    __shared__ int64_t red[256 / WAVE];
    const int32_t ns = mc[2], total = mc[2] + mc[3];
    int64_t cc = 0;
    FOR_THREAD(x, total) {
        vid_t v = list_at(moves, ns, nmax, x);
        const int2* tab = conn + conn_off[v];
        int32_t sz = conn_size[v];
        cc += (int64_t)lookup(tab, sz, part[v]) - lookup(tab, sz, dest[v]);
    }
    cc = block_reduce_sum<int64_t, 256>(cc, red);
    if (threadIdx.x == 0 && cc) atomicAdd((unsigned long long*)&acc[1], (unsigned long long)cc);
    if (blockIdx.x == 0) {
        int64_t mx = INT64_MIN;
        for (int t = threadIdx.x; t < k; t += blockDim.x) mx = sizes[t] > mx ? sizes[t] : mx;
        mx = block_reduce_max<int64_t, 256>(mx, red);
        if (threadIdx.x == 0) acc[2] = mx;
    }
}

// ---------------------------------------------------------------- rebalancing
__global__ void k_undersized(const int64_t* sizes, int32_t k, int64_t max_dest, part_t* undersized, int32_t* misc) {
    __shared__ int32_t ws[256 / WAVE + 1];
    int32_t base = 0;
    for (int32_t p0 = 0; p0 < k; p0 += 256) {
        int32_t p = p0 + threadIdx.x;
        int32_t f = (p < k && sizes[p] < max_dest) ? 1 : 0;
        int32_t tot;
        int32_t ex = block_excl_scan<int32_t, 256>(f, tot, ws);
        if (f) undersized[base + ex] = p;
        base += tot;
    }
    if (threadIdx.x == 0) misc[0] = base;
}

// jet_refiner.hpp:592-643: destination and bucket of each vertex of an
// oversized part.
__global__ void k_weak_select(DGraph g, const part_t* part, const int64_t* sizes, int64_t size_max, int32_t opt,
                              int64_t max_dest, const part_t* undersized, const int32_t* misc,
                              const eoff_t* conn_off, const int2* conn, const int32_t* conn_size, int32_t sections,
                              part_t* dest, int32_t* bid, int32_t* flags) {
    const int32_t tu = misc[0];
    FOR_THREAD(i, g.n) {
        part_t p = part[i];
        part_t d = p;
        int32_t f = 0;
        int64_t sp = sizes[p];
        if (sp > size_max && (double)g.vw[i] < 1.5 * (double)(sp - opt)) {
            part_t best = p;
            int32_t gain = 0, p_gain = 0;
            const int2* tab = conn + conn_off[i];
            int32_t sz = conn_size[i];
            for (int32_t s = 0; s < sz; s++) {
                int2 e = tab[s];
                if (e.x > NULL_PART && sizes[e.x] < max_dest) {
                    if (e.y > gain) {
                        best = e.x;
                        gain = e.y;
                    } else if (e.y == gain && gain > 0 && e.x < best) {
                        best = e.x;
                    }
                }
                if (e.x == p) p_gain = e.y;
            }
            int64_t sv = 0;
            if (gain > 0) {
                d = best;
                sv = (int64_t)gain - p_gain;
            } else if (tu > 0) {
                d = undersized[i % tu];
                sv = -(int64_t)p_gain;
            }
            if (d != p) {
                int32_t gt = gain_bucket(sv, g.vw[i]);
                bid[i] = (MAX_BUCKETS * p + gt) * sections + (int32_t)(i % sections);
                f = 1;
            }
        }
        dest[i] = d;
        flags[i] = f;
    }
}

__global__ void k_strong_prep(const int64_t* sizes, int32_t k, int64_t max_dest, int64_t* acc) {
    __shared__ int64_t red[256 / WAVE];
    int64_t mx = INT64_MIN;
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
        int64_t sz = sizes[t];
        if (sz < max_dest && max_dest - sz > mx) mx = max_dest - sz;
    }
    mx = block_reduce_max<int64_t, 256>(mx, red);
    if (threadIdx.x == 0) acc[3] = mx;
}

// jet_refiner.hpp:449-481.
__global__ void k_strong_select(DGraph g, const part_t* part, const int64_t* sizes, int64_t size_max, int32_t opt,
                                int64_t max_dest, const int64_t* acc, const eoff_t* conn_off, const int2* conn,
                                const int32_t* conn_size, int32_t sections, int32_t* bid, int32_t* flags) {
    const int64_t max_vwgt = acc[3];
    FOR_THREAD(i, g.n) {
        part_t p = part[i];
        int32_t f = 0;
        int64_t sp = sizes[p];
        int64_t w = g.vw[i];
        bool vw_ok = max_vwgt > 0 && w <= 2 * max_vwgt;
        if (sp > size_max && vw_ok && w < 2 * (sp - opt)) {
            uint64_t tk = 0, tg = 0;
            int64_t p_gain = 0;
            const int2* tab = conn + conn_off[i];
            int32_t sz = conn_size[i];
            for (int32_t s = 0; s < sz; s++) {
                int2 e = tab[s];
                if (e.x == p) {
                    p_gain = e.y;
                } else if (e.x > NULL_PART && sizes[e.x] < max_dest) {
                    tg += (uint64_t)e.y;
                    tk += 1;
                }
            }
            if (tk == 0) tk = 1;
            int64_t gain = (int64_t)(tg / tk) - p_gain;
            int64_t vwgt = w < sp - size_max ? w : sp - size_max;
            int32_t gt = gain_bucket(gain, vwgt);
            if (gt < MAX_BUCKETS) {
                bid[i] = (MAX_BUCKETS * p + gt) * sections + (int32_t)(i % sections) + 1;
                f = 1;
            }
        }
        flags[i] = f;
    }
}

__global__ void k_gather_cands(vid_t n, const int32_t* pos, const int32_t* bid, uint32_t* skey, uint32_t* sval) {
    FOR_THREAD(i, n) {
        int32_t p0 = pos[i];
        if (pos[i + 1] != p0) {
            skey[p0] = (uint32_t)bid[i];
            sval[p0] = (uint32_t)i;
        }
    }
}

__global__ void k_sorted_weights(int32_t nc, const uint32_t* sval, const wgt_t* vw, int64_t* w) {
    FOR_THREAD(t, nc) w[t] = vw[sval[t]];
}

__device__ inline int32_t lower_bound_u32(const uint32_t* a, int32_t n, uint32_t key) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_part_begin(const uint32_t* skey, int32_t nc, int32_t k, int64_t width, int32_t* pb) {
    FOR_THREAD(p, (int64_t)k + 1) pb[p] = lower_bound_u32(skey, nc, (uint32_t)((int64_t)p * width));
}

// get_evictions (jet_refiner.hpp:359-412): score = weight of this part's
// candidates ranked before the vertex (bucket order, then vertex id).
template <bool ADJUST>
__global__ void k_evict(DGraph g, const uint32_t* sval, int32_t nc, const int64_t* prefix, const int32_t* pb,
                        const part_t* part, const int64_t* sizes, int64_t size_max, int64_t* evict_adjust,
                        int64_t* score, vid_t* moves, int32_t* mcounters, vid_t nmax) {
    FOR_THREAD(t, nc) {
        vid_t v = (vid_t)sval[t];
        part_t p = part[v];
        int64_t sc = prefix[t] - prefix[pb[p]];
        int64_t limit = sizes[p] - size_max;
        if (sc < limit) {
            if (ADJUST) {
                if (sc + g.vw[v] >= limit) evict_adjust[p] = sc + g.vw[v];
                score[v] = prefix[t];
            }
            push_list(moves, mcounters, 2, nmax, v, g.rowptr[v + 1] - g.rowptr[v]);
        }
    }
}

// Cookie cutter (jet_refiner.hpp:491-518): per-part eviction totals become a
// global eviction order; receivers get consecutive chunks of capacity.
__global__ void k_cookie(int32_t k, const int64_t* prefix, const int32_t* pb, const int64_t* sizes, int64_t max_dest,
                         int64_t* evict_adjust, int64_t* evict_start, int64_t* acc) {
    if (threadIdx.x != 0) return;
    int64_t run = 0;
    for (int32_t p = 0; p < k; p++) {
        int64_t add = evict_adjust[p];
        if (add == 0) add = prefix[pb[p + 1]] - prefix[pb[p]];
        evict_adjust[p] = prefix[pb[p]] - run;
        run += add;
    }
    acc[3] = run;  // re-draw counter starts after every eviction
    int64_t run2 = 0;
    evict_start[0] = 0;
    for (int32_t p = 0; p < k; p++) {
        if (max_dest > sizes[p]) run2 += max_dest - sizes[p];
        evict_start[p + 1] = run2;
    }
}

// Last index q in [lo, k] with evict_start[q] <= s (evict_start is non-decreasing).
__device__ inline int32_t chunk_of(const int64_t* es, int32_t lo, int32_t k, int64_t s) {
    int32_t a = lo, b = k + 1;  // first index with es > s lies in [a, b]
    while (a < b) {
        int32_t mid = (a + b) >> 1;
        if (es[mid] <= s) a = mid + 1;
        else b = mid;
    }
    return a - 1;
}

// jet_refiner.hpp:519-544 without the re-draws, which are queued.
__global__ void k_strong_dest(const vid_t* moves, const int32_t* mc, vid_t nmax, const part_t* part, const wgt_t* vw,
                              const int64_t* evict_adjust, const int64_t* evict_start, int32_t k, int64_t* score,
                              part_t* dest, vid_t* redraw, int32_t* rcount, int32_t rmax) {
    const int32_t ns = mc[2], total = mc[2] + mc[3];
    FOR_THREAD(x, total) {
        vid_t v = list_at(moves, ns, nmax, x);
        part_t pv = part[v];
        int64_t s = score[v] - evict_adjust[pv];
        score[v] = s;
        int32_t q = chunk_of(evict_start, 0, k, s);
        if (q < k && vw[v] / 2 <= evict_start[q + 1] - s) {
            dest[v] = q;
        } else if (q < k) {
            int32_t r = atomicAdd(rcount, 1);
            if (r < rmax) redraw[r] = v;
            dest[v] = pv;
        } else {
            dest[v] = pv;
        }
    }
}

// Re-draws in increasing vertex id, one after another (the counter starts at
// the total eviction weight).
constexpr int RD_MAX = 8192;
__global__ void __launch_bounds__(256) k_strong_redraw(vid_t* redraw, const int32_t* rcount, const part_t* part,
                                                       const wgt_t* vw, const int64_t* evict_start, int32_t k,
                                                       const int64_t* score, part_t* dest, int64_t* acc, int32_t* err) {
    __shared__ vid_t ids[RD_MAX];
    int32_t r = rcount[0];
    if (r > RD_MAX) {
        if (threadIdx.x == 0) atomicOr(err, 4);
        return;
    }
    uint32_t cap = 1;
    while ((int32_t)cap < r) cap <<= 1;
    for (uint32_t t = threadIdx.x; t < cap; t += blockDim.x) ids[t] = (int32_t)t < r ? redraw[t] : ORD_MAX;
    for (uint32_t size = 2; size <= cap; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < cap / 2; t += blockDim.x) {
                uint32_t lo = 2 * t - (t & (stride - 1));
                uint32_t hi = lo + stride;
                bool up = (lo & size) == 0;
                vid_t a = ids[lo], b = ids[hi];
                if ((a > b) == up) {
                    ids[lo] = b;
                    ids[hi] = a;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int64_t counter = acc[3];
    for (int32_t x = 0; x < r; x++) {
        vid_t v = ids[x];
        int64_t s = score[v];
        int32_t p = 0;
        bool done = false;
        while (p < k) {
            p = chunk_of(evict_start, p, k, s);
            if (p < k && vw[v] / 2 <= evict_start[p + 1] - s) {
                dest[v] = p;
                done = true;
                break;
            } else if (p < k) {
                s = counter;
                counter += vw[v];
            }
        }
        if (!done) dest[v] = part[v];
    }
    acc[3] = counter;
}

__global__ void k_collect_moves(const vid_t* moves, const int32_t* mc, vid_t nmax, const part_t* dest, vid_t* out_v,
                                part_t* out_d) {
    const int32_t total = mc[2] + mc[3];
    FOR_THREAD(x, total) {
        vid_t v = list_at(moves, mc[2], nmax, x);
        out_v[x] = v;
        out_d[x] = dest[v];
    }
}

// ---------------------------------------------------------------- host side
static inline int32_t sections_for(vid_t n, int32_t k) {
    int32_t sections = MAX_SECTIONS;
    int32_t section_size = (n + sections * k) / (sections * k);
    if (section_size < 4096) {
        section_size = 4096;
        sections = (n + section_size * k) / (section_size * k);
    }
    return sections;
}

Refiner::Refiner(Arena& a, hipStream_t s, vid_t n, int64_t cap, const RefineParams& p) : a_(a), s_(s), p_(p), nmax_(n) {
    if (p.k < 1 || p.k > MAX_K) throw std::invalid_argument("k out of range");
    int64_t N = std::max<int64_t>(n, 1);
    conn_off_ = a.alloc<eoff_t>(N + 1);
    conn_ = a.alloc<int2>(std::max<int64_t>(cap, 1));
    conn_size_ = a.alloc<int32_t>(N);
    dest_cache_ = a.alloc<part_t>(N);
    lock_ = a.alloc<int32_t>(N);
    gain_p_ = a.alloc<int32_t>(N);
    pregain_ = a.alloc<int32_t>(N);
    dest_ = a.alloc<part_t>(N);
    save_ = a.alloc<int32_t>(N);
    bid_ = a.alloc<int32_t>(N);
    flags_ = a.alloc<int32_t>(N + 1);
    cand_ = a.alloc<vid_t>(N);
    moves_ = a.alloc<vid_t>(N);
    swap_bit_ = a.alloc<int32_t>(N);
    skey_ = a.alloc<uint32_t>(N);
    sval_ = a.alloc<uint32_t>(N);
    ktmp_ = a.alloc<uint32_t>(N);
    vtmp_ = a.alloc<uint32_t>(N);
    prefix_ = a.alloc<int64_t>(N + 1);
    score_ = a.alloc<int64_t>(N);
    cur_part_ = a.alloc<part_t>(N);
    tier_mid_ = a.alloc<vid_t>(N);
    tier_big_ = a.alloc<vid_t>(N);
    counters_ = a.alloc<int32_t>(16);
    acc_ = a.alloc<int64_t>(8);
    curr_sizes_ = a.alloc<int64_t>(p.k);
    best_sizes_ = a.alloc<int64_t>(p.k);
    evict_adjust_ = a.alloc<int64_t>(p.k);
    evict_start_ = a.alloc<int64_t>(p.k + 1);
    part_begin_ = a.alloc<int32_t>(p.k + 1);
    undersized_ = a.alloc<part_t>(p.k);
    misc_ = a.alloc<int32_t>(8);
    redraw_ = a.alloc<vid_t>(RD_MAX);
    fill_i32(swap_bit_, 0, N, s);
    fill_i32(counters_, 0, 16, s);
}

void Refiner::build_tiers(const Level& L) {
    DGraph g = L.view();
    fill_i32(counters_ + 8, 0, 2, s_);
    k_conn_tiers<<<grid_for(g.n, 256, 8192), 256, 0, s_>>>(g, p_.k, tier_mid_, tier_big_, counters_ + 8);
    JET_LAUNCH_CHECK();
    int32_t c[2];
    JET_HIP(hipMemcpyAsync(c, counters_ + 8, sizeof(c), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    ntier_mid_ = c[0];
    ntier_big_ = c[1];
}

void Refiner::init_conn(const Level& L, const part_t* part, const int32_t* mark) {
    DGraph g = L.view();
    int32_t* mk = const_cast<int32_t*>(mark);
    if (!mark) {
        ArenaScope sc(a_);
        int32_t* caps = a_.alloc<int32_t>((int64_t)g.n + 1);
        k_conn_caps<<<grid_for(g.n, 256, 8192), 256, 0, s_>>>(g, p_.k, caps);
        JET_LAUNCH_CHECK();
        excl_scan_i32_i64(caps, conn_off_, g.n, a_, s_);
        fill_i32(dest_cache_, NULL_PART, g.n, s_);
        fill_i32(lock_, 0, g.n, s_);
    }
    k_conn_build_small<<<grid_for(g.n, SB_BLOCK, 8192), SB_BLOCK, 0, s_>>>(g, part, p_.k, conn_off_, conn_, conn_size_,
                                                                           mk, dest_cache_);
    JET_LAUNCH_CHECK();
    if (ntier_mid_ > 0) {
        k_conn_build_wave<<<grid_for(ntier_mid_, WB_WAVES, 8192), WB_WAVES * WAVE, 0, s_>>>(
            g, part, p_.k, conn_off_, conn_, conn_size_, mk, dest_cache_, tier_mid_, ntier_mid_);
        JET_LAUNCH_CHECK();
    }
    if (ntier_big_ > 0) {
        k_conn_build_block<<<std::min(ntier_big_, 2048), BB_BLOCK, 0, s_>>>(g, part, p_.k, conn_off_, conn_, conn_size_,
                                                                            mk, dest_cache_, tier_big_, ntier_big_);
        JET_LAUNCH_CHECK();
    }
}

Refiner::Moves Refiner::jet_lp(const Level& L, const part_t* part, double fr) {
    DGraph g = L.view();
    fill_i32(counters_, 0, 4, s_);
    int nb = grid_for(g.n, 256, 8192);
    k_lp_select<<<nb, 256, 0, s_>>>(g, part, conn_off_, conn_, conn_size_, dest_cache_, gain_p_, lock_, fr, dest_,
                                    pregain_, cand_, counters_, nmax_);
    JET_LAUNCH_CHECK();
    k_afterburner_small<<<nb, 256, 0, s_>>>(g, cand_, counters_, part, dest_, pregain_, lock_, moves_, counters_, nmax_);
    k_afterburner_big<<<grid_for((int64_t)g.n * WAVE, 256, 4096), 256, 0, s_>>>(g, cand_, counters_, part, dest_,
                                                                               pregain_, lock_, moves_, counters_, nmax_);
    JET_LAUNCH_CHECK();
    int32_t c[4];
    JET_HIP(hipMemcpyAsync(c, counters_, sizeof(c), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    return Moves{c[2], c[3]};
}

// Candidates flagged in flags_ with keys in bid_: ordered compaction, stable
// sort by key, prefix of weights in sorted order, part boundaries.
void Refiner::sort_candidates(const Level& L, int32_t ncand, int key_bits) {
    DGraph g = L.view();
    k_gather_cands<<<grid_for(g.n, 256, 8192), 256, 0, s_>>>(g.n, flags_, bid_, skey_, sval_);
    JET_LAUNCH_CHECK();
    radix_sort_u32(skey_, sval_, ncand, key_bits, ktmp_, vtmp_, a_, s_);
    if (ncand > 0) {
        k_sorted_weights<<<grid_for(ncand, 256, 8192), 256, 0, s_>>>(ncand, sval_, g.vw, prefix_);
        JET_LAUNCH_CHECK();
    }
    excl_scan_i64_i64(prefix_, prefix_, ncand, a_, s_);
}

Refiner::Moves Refiner::rebalance_weak(const Level& L, const part_t* part, const int64_t* sizes) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    int32_t sections = sections_for(g.n, k);
    int64_t width = (int64_t)MAX_BUCKETS * sections;
    int64_t smax = size_max_;
    int64_t max_dest = (int64_t)(smax * 0.99);
    if (max_dest < smax - 100) max_dest = smax - 100;
    fill_i32(counters_, 0, 4, s_);
    k_undersized<<<1, 256, 0, s_>>>(sizes, k, max_dest, undersized_, misc_);
    int nb = grid_for(g.n, 256, 8192);
    k_weak_select<<<nb, 256, 0, s_>>>(g, part, sizes, smax, opt_, max_dest, undersized_, misc_, conn_off_, conn_,
                                      conn_size_, sections, dest_, bid_, flags_);
    JET_LAUNCH_CHECK();
    excl_scan_i32_i32(flags_, flags_, g.n, a_, s_);
    int32_t ncand = read_scalar(flags_ + g.n, s_);
    int64_t maxkey = (int64_t)k * width + sections;
    int bits = 1;
    while ((int64_t(1) << bits) <= maxkey) bits++;
    sort_candidates(L, ncand, bits);
    k_part_begin<<<grid_for((int64_t)k + 1, 256, 64), 256, 0, s_>>>(skey_, ncand, k, width, part_begin_);
    if (ncand > 0) {
        k_evict<false><<<grid_for(ncand, 256, 8192), 256, 0, s_>>>(g, sval_, ncand, prefix_, part_begin_, part, sizes,
                                                                   smax, evict_adjust_, score_, moves_, counters_,
                                                                   nmax_);
    }
    JET_LAUNCH_CHECK();
    int32_t c[4];
    JET_HIP(hipMemcpyAsync(c, counters_, sizeof(c), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    return Moves{c[2], c[3]};
}

Refiner::Moves Refiner::rebalance_strong(const Level& L, const part_t* part, const int64_t* sizes) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    int32_t sections = sections_for(g.n, k);
    int64_t width = (int64_t)MAX_BUCKETS * sections;
    int64_t smax = size_max_;
    int64_t max_dest = std::max<int64_t>((int64_t)opt_ + 1, (int64_t)(smax * 0.99));
    fill_i32(counters_, 0, 5, s_);
    k_strong_prep<<<1, 256, 0, s_>>>(sizes, k, max_dest, acc_);
    int nb = grid_for(g.n, 256, 8192);
    k_strong_select<<<nb, 256, 0, s_>>>(g, part, sizes, smax, opt_, max_dest, acc_, conn_off_, conn_, conn_size_,
                                        sections, bid_, flags_);
    JET_LAUNCH_CHECK();
    excl_scan_i32_i32(flags_, flags_, g.n, a_, s_);
    int32_t ncand = read_scalar(flags_ + g.n, s_);
    int64_t maxkey = (int64_t)k * width + 1;
    int bits = 1;
    while ((int64_t(1) << bits) <= maxkey) bits++;
    sort_candidates(L, ncand, bits);
    k_part_begin<<<grid_for((int64_t)k + 1, 256, 64), 256, 0, s_>>>(skey_, ncand, k, width, part_begin_);
    fill_i64(evict_adjust_, 0, k, s_);
    if (ncand > 0) {
        k_evict<true><<<grid_for(ncand, 256, 8192), 256, 0, s_>>>(g, sval_, ncand, prefix_, part_begin_, part, sizes,
                                                                  smax, evict_adjust_, score_, moves_, counters_, nmax_);
    }
    k_cookie<<<1, 64, 0, s_>>>(k, prefix_, part_begin_, sizes, max_dest, evict_adjust_, evict_start_, acc_);
    k_strong_dest<<<nb, 256, 0, s_>>>(moves_, counters_, nmax_, part, g.vw, evict_adjust_, evict_start_, k, score_,
                                      dest_, redraw_, counters_ + 4, RD_MAX);
    k_strong_redraw<<<1, 256, 0, s_>>>(redraw_, counters_ + 4, part, g.vw, evict_start_, k, score_, dest_, acc_,
                                       misc_ + 1);
    JET_LAUNCH_CHECK();
    int32_t c[4];
    JET_HIP(hipMemcpyAsync(c, counters_, sizeof(c), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    return Moves{c[2], c[3]};
}

void Refiner::perform_moves(const Level& L, part_t* part, Moves mv, int64_t* sizes, int64_t& cut_change,
                            int64_t& max_size) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    fill_i64(acc_, 0, 3, s_);
    int32_t total = mv.total();
    int nb = grid_for(std::max(total, 1), 256, 8192);
    k_moves_apply<<<nb, 256, 0, s_>>>(moves_, counters_, nmax_, g.vw, part, dest_, dest_cache_, conn_off_, conn_,
                                      conn_size_, reinterpret_cast<unsigned long long*>(sizes), k, acc_);
    JET_LAUNCH_CHECK();
    if (total > g.n / 10) {
        k_mark_adjacent<<<nb, 256, 0, s_>>>(g, moves_, counters_, nmax_, swap_bit_);
        JET_LAUNCH_CHECK();
        init_conn(L, part, swap_bit_);
    } else if (total > 0) {
        int nbs = grid_for(std::max(mv.nsmall, 1), 256, 8192);
        int nbb = grid_for((int64_t)std::max(mv.nbig, 1) * WAVE, 256, 4096);
        k_update_small<false><<<nbs, 256, 0, s_>>>(g, moves_, counters_, nmax_, part, dest_, conn_off_, conn_, conn_size_,
                                                   dest_cache_, k, misc_ + 1);
        k_update_big<false><<<nbb, 256, 0, s_>>>(g, moves_, counters_, nmax_, part, dest_, conn_off_, conn_, conn_size_,
                                                 dest_cache_, k, misc_ + 1);
        k_update_small<true><<<nbs, 256, 0, s_>>>(g, moves_, counters_, nmax_, part, dest_, conn_off_, conn_, conn_size_,
                                                  dest_cache_, k, misc_ + 1);
        k_update_big<true><<<nbb, 256, 0, s_>>>(g, moves_, counters_, nmax_, part, dest_, conn_off_, conn_, conn_size_,
                                                dest_cache_, k, misc_ + 1);
        JET_LAUNCH_CHECK();
    }
    k_moves_finish<<<nb, 256, 0, s_>>>(moves_, counters_, nmax_, part, dest_, conn_off_, conn_, conn_size_, sizes, k,
                                       acc_);
    JET_LAUNCH_CHECK();
    int64_t h[3];
    JET_HIP(hipMemcpyAsync(h, acc_, sizeof(h), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    cut_change = h[0] + h[1];
    max_size = h[2];
}

LevelRecord Refiner::refine(const Level& L, part_t* best_part) {
    auto t0 = std::chrono::steady_clock::now();
    DGraph g = L.view();
    const int32_t k = p_.k;
    fill_i32(misc_, 0, 8, s_);
    if (!init_) {
        cut_ = total_cut2(g, best_part, a_, s_);
        part_sizes(g, best_part, k, best_sizes_, s_);
        std::vector<int64_t> hs(k);
        JET_HIP(hipMemcpyAsync(hs.data(), best_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToHost, s_));
        JET_HIP(hipStreamSynchronize(s_));
        total_size_ = 0;
        for (int64_t v : hs) total_size_ += v;
    }
    opt_ = (int32_t)(((int32_t)total_size_ + k - 1) / k);
    size_max_ = (int32_t)(opt_ * p_.imb);
    if (!init_) {
        init_ = true;
        std::vector<int64_t> hs(k);
        JET_HIP(hipMemcpyAsync(hs.data(), best_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToHost, s_));
        JET_HIP(hipStreamSynchronize(s_));
        int64_t mx = 0;
        for (int64_t v : hs) mx = std::max(mx, v);
        total_imb_ = mx > opt_ ? mx - opt_ : 0;
    }
    int64_t curr_cut = cut_, curr_imb = total_imb_;
    JET_HIP(hipMemcpyAsync(curr_sizes_, best_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToDevice, s_));
    const int64_t imb_max = (int64_t)size_max_ - opt_;
    copy_i32(cur_part_, best_part, g.n, s_);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    std::vector<double> temps;
    if (p_.ultra) {
        for (double t = 0.85; t > 0; t -= 0.05) temps.push_back(t);
    } else if (L.uniform) {
        temps.push_back(0.25);
    } else {
        temps.push_back(0.75);
    }
    LevelRecord rec;
    int balance_counter = 0;
    for (double fr : temps) {
        int count = 0;
        while (count++ <= 11) {
            rec.iterations++;
            Moves mv;
            if (curr_imb <= imb_max) {
                mv = jet_lp(L, cur_part_, fr);
                balance_counter = 0;
                rec.lp_iterations++;
            } else {
                if (balance_counter < 2) mv = rebalance_weak(L, cur_part_, curr_sizes_);
                else mv = rebalance_strong(L, cur_part_, curr_sizes_);
                balance_counter++;
            }
            int64_t cc = 0, mx = 0;
            perform_moves(L, cur_part_, mv, curr_sizes_, cc, mx);
            curr_imb = mx > opt_ ? mx - opt_ : 0;
            curr_cut -= cc;
            bool take = false;
            if (total_imb_ > imb_max && curr_imb < total_imb_) {
                take = true;
                count = 0;
            } else if (curr_cut < cut_ && (curr_imb <= imb_max || curr_imb <= total_imb_)) {
                if ((double)curr_cut < p_.tol * (double)cut_) count = 0;
                take = true;
            }
            if (take) {
                cut_ = curr_cut;
                total_imb_ = curr_imb;
                JET_HIP(hipMemcpyAsync(best_sizes_, curr_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToDevice, s_));
                copy_i32(best_part, cur_part_, g.n, s_);
            }
        }
    }
    int32_t err = read_scalar(misc_ + 1, s_);
    if (err) throw DeviceError("connectivity table invariant violated (code " + std::to_string(err) + ")");
    rec.cut2 = cut_;
    rec.total_imb = total_imb_;
    rec.n = g.n;
    rec.m = g.m;
    rec.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rec;
}

void Refiner::single_step(const Level& L, const part_t* part, int kind, double fr, std::vector<int32_t>& moves,
                          std::vector<int32_t>& dest) {
    DGraph g = L.view();
    const int32_t k = p_.k;
    fill_i32(misc_, 0, 8, s_);
    copy_i32(cur_part_, part, g.n, s_);
    part_sizes(g, cur_part_, k, curr_sizes_, s_);
    std::vector<int64_t> hs(k);
    JET_HIP(hipMemcpyAsync(hs.data(), curr_sizes_, sizeof(int64_t) * k, hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    int64_t tot = 0;
    for (int64_t v : hs) tot += v;
    opt_ = (int32_t)(((int32_t)tot + k - 1) / k);
    size_max_ = (int32_t)(opt_ * p_.imb);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    Moves mv;
    if (kind == 0) mv = jet_lp(L, cur_part_, fr);
    else if (kind == 1) mv = rebalance_weak(L, cur_part_, curr_sizes_);
    else mv = rebalance_strong(L, cur_part_, curr_sizes_);
    int32_t total = mv.total();
    ArenaScope sc(a_);
    vid_t* ov = a_.alloc<vid_t>(std::max(total, 1));
    part_t* od = a_.alloc<part_t>(std::max(total, 1));
    if (total > 0) {
        k_collect_moves<<<grid_for(total, 256, 8192), 256, 0, s_>>>(moves_, counters_, nmax_, dest_, ov, od);
        JET_LAUNCH_CHECK();
    }
    std::vector<int32_t> v(total), d(total);
    if (total > 0) {
        JET_HIP(hipMemcpyAsync(v.data(), ov, sizeof(int32_t) * total, hipMemcpyDeviceToHost, s_));
        JET_HIP(hipMemcpyAsync(d.data(), od, sizeof(int32_t) * total, hipMemcpyDeviceToHost, s_));
    }
    JET_HIP(hipStreamSynchronize(s_));
    int32_t err = read_scalar(misc_ + 1, s_);
    if (err) throw DeviceError("refinement invariant violated (code " + std::to_string(err) + ")");
    std::vector<int32_t> idx(total);
    for (int32_t t = 0; t < total; t++) idx[t] = t;
    std::sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return v[x] < v[y]; });
    moves.resize(total);
    dest.resize(total);
    for (int32_t t = 0; t < total; t++) {
        moves[t] = v[idx[t]];
        dest[t] = d[idx[t]];
    }
}

void Refiner::conn_tables(const Level& L, const part_t* part, std::vector<int64_t>& off,
                          std::vector<int32_t>& slot_part, std::vector<int32_t>& slot_val, std::vector<int32_t>& size) {
    DGraph g = L.view();
    copy_i32(cur_part_, part, g.n, s_);
    build_tiers(L);
    init_conn(L, cur_part_, nullptr);
    off.resize(g.n + 1);
    size.resize(g.n);
    JET_HIP(hipMemcpyAsync(off.data(), conn_off_, sizeof(int64_t) * (g.n + 1), hipMemcpyDeviceToHost, s_));
    JET_HIP(hipMemcpyAsync(size.data(), conn_size_, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    int64_t C = off[g.n];
    std::vector<int2> slots(std::max<int64_t>(C, 1));
    if (C > 0) JET_HIP(hipMemcpyAsync(slots.data(), conn_, sizeof(int2) * C, hipMemcpyDeviceToHost, s_));
    JET_HIP(hipStreamSynchronize(s_));
    slot_part.resize(C);
    slot_val.resize(C);
    for (int64_t c = 0; c < C; c++) {
        slot_part[c] = slots[c].x;
        slot_val[c] = slots[c].y;
    }
}

}  // namespace jet
