// Multilevel partition orchestration and the C-ABI (include/jet_capi.h).
//
// partition_once() follows partitioner::partition (partitioner.hpp:61-125):
// heuristic choice and coarsening cutoff, the coarsening loop of
// contracter::generate_coarse_graphs (contract.hpp:494-533), the initial
// partition of the coarsest graph on the host (initial_partition.hpp:92-114),
// then uncoarsener::multilevel_jet (uncoarsen.hpp:88-125): refine each level
// from coarsest to finest and project onto the next finer one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "arena.h"
#include "coarsen.h"
#include "engine.h"
#include "hierarchy.h"
#include "contract.h"
#include "jet_capi.h"
#include "misc.h"
#include "prims.h"
#include "prof.h"
#include "refine.h"
#include "tuning.h"

#include "initpart.h"

namespace jet {

struct InvalidInput : std::runtime_error {
    explicit InvalidInput(const std::string& s) : std::runtime_error(s) {}
};
struct InitFailure : std::runtime_error {
    explicit InitFailure(const std::string& s) : std::runtime_error(s) {}
};

Context* context_create(int device) {
    int cnt = 0;
    hipError_t e = hipGetDeviceCount(&cnt);
    if (e != hipSuccess || cnt < 1) {
        (void)hipGetLastError();
        throw DeviceError("no HIP device available");
    }
    if (device < 0 || device >= cnt) throw InvalidInput("device index out of range");
    JET_HIP(hipSetDevice(device));
    {
        // the device's tuning table, once per process
        hipDeviceProp_t prop{};
        JET_HIP(hipGetDeviceProperties(&prop, device));
        std::string arch = prop.gcnArchName;  // e.g. "gfx950:sramecc+:xnack-"
        tuning_init_once(arch.substr(0, arch.find(':')));
    }
    Context* c = new Context();
    c->device = device;
    JET_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    return c;
}

void context_destroy(Context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->persist.free_all();
    c->scratch.free_all();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

static std::mutex g_ctx_mu;
static Context* g_ctx = nullptr;

// The process-wide context of the plain entry points (device 0).
static Context& context() {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx) g_ctx = context_create(0);
    JET_HIP(hipSetDevice(g_ctx->device));
    return *g_ctx;
}

// Rewinds both arenas when a call ends (chunks stay cached).
struct CallScope {
    Context& c;
    std::lock_guard<std::mutex> lk;
    explicit CallScope(Context& ctx) : c(ctx), lk(ctx.mu) {
        c.persist.reset();
        c.scratch.reset();
    }
    ~CallScope() {
        (void)hipStreamSynchronize(c.stream);
        c.persist.reset();
        c.scratch.reset();
    }
};

static thread_local std::string g_err;

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

template <class T>
static T* to_device(Context& C, const T* src, int64_t count, int space) {
    if (space == JET_PTR_DEVICE) return const_cast<T*>(src);
    T* d = C.persist.alloc<T>(count);
    if (count > 0) JET_HIP(hipMemcpyAsync(d, src, sizeof(T) * count, hipMemcpyHostToDevice, C.stream));
    return d;
}

template <class T>
static void from_device(Context& C, T* dst, const T* src, int64_t count, int space) {
    if (count <= 0) return;
    if (space == JET_PTR_DEVICE) {
        if (dst != src) JET_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToDevice, C.stream));
    } else {
        JET_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, C.stream));
    }
    JET_HIP(hipStreamSynchronize(C.stream));
}

template <class T>
static T read_one(Context& C, const T* p, int space) {
    if (space == JET_PTR_HOST) return *p;
    return read_scalar(p, C.stream);
}

static void check_k(int32_t k) {
    if (k < 1 || k > JET_MAX_PARTS) throw InvalidInput("k must be in [1, 8192]");
}

// Part ids index per-part arrays on the device: reject ids outside [0, k).
static void check_parts(Context& C, int64_t n, const part_t* dev_part, int32_t k) {
    if (!parts_in_range(n, dev_part, k, C.scratch, C.stream)) throw InvalidInput("part ids must be in [0, k)");
}

static void check_space(int space) {
    if (space != JET_PTR_HOST && space != JET_PTR_DEVICE) throw InvalidInput("ptr_space must be JET_PTR_HOST or JET_PTR_DEVICE");
}

// Stages a CSR graph on the device (copies host arrays, widens int32 offsets,
// narrows int64 weights, materialises unit weights) and validates it.
static Level stage_graph(Context& C, int64_t n, const int64_t* rp64, const int32_t* rp32, const int32_t* col,
                         const int32_t* ew32, const int64_t* ew64, const int32_t* vw32, const int64_t* vw64, int space,
                         bool uniform, bool validate = true, const int64_t* col64 = nullptr) {
    check_space(space);
    if (n < 1 || n > (int64_t)ORD_MAX - 2) throw InvalidInput("n must be in [1, 2^31 - 2)");
    if ((!rp64 && !rp32) || (!col && !col64)) throw InvalidInput("row offsets and columns are required");
    hipStream_t s = C.stream;
    Level L;
    L.n = (vid_t)n;
    L.uniform = uniform;
    if (rp64) {
        L.m = read_one(C, rp64 + n, space);
        int64_t first = read_one(C, rp64, space);
        if (first != 0) throw InvalidInput("row_offsets[0] must be 0");
        L.rowptr = to_device(C, rp64, n + 1, space);
    } else {
        L.m = read_one(C, rp32 + n, space);
        int32_t first = read_one(C, rp32, space);
        if (first != 0) throw InvalidInput("row_offsets[0] must be 0");
        const int32_t* d32 = to_device(C, rp32, n + 1, space);
        L.rowptr = C.persist.alloc<eoff_t>(n + 1);
        widen_offsets(n + 1, d32, L.rowptr, s);
    }
    if (L.m < 0) throw InvalidInput("row_offsets[n] must be >= 0");
    if (col64) {
        const int64_t* d = to_device(C, col64, L.m, space);
        L.col = C.persist.alloc<vid_t>(L.m);
        if (narrow_columns(L.m, d, n, L.col, C.scratch, s)) throw InvalidInput("invalid graph: column out of range;");
    } else {
        L.col = to_device(C, col, L.m, space);
    }
    if (ew64) {
        const int64_t* d = to_device(C, ew64, L.m, space);
        L.ew = C.persist.alloc<wgt_t>(L.m);
        if (narrow_weights(L.m, d, L.ew, C.scratch, s)) throw InvalidInput("edge weights must be in [1, 2^31)");
    } else if (ew32) {
        L.ew = to_device(C, ew32, L.m, space);
    } else {
        L.ew = C.persist.alloc<wgt_t>(L.m);
        fill_i32(L.ew, 1, L.m, s);
    }
    if (vw64) {
        const int64_t* d = to_device(C, vw64, n, space);
        L.vw = C.persist.alloc<wgt_t>(n);
        if (narrow_weights(n, d, L.vw, C.scratch, s)) throw InvalidInput("vertex weights must be in [1, 2^31)");
    } else if (vw32) {
        L.vw = to_device(C, vw32, n, space);
    } else {
        L.vw = C.persist.alloc<wgt_t>(n);
        fill_i32(L.vw, 1, n, s);
    }
    if (validate) {
        int32_t e = validate_graph(L.n, L.rowptr, L.col, L.ew, L.vw, C.scratch, s);
        if (e) {
            std::string msg = "invalid graph:";
            if (e & 1) msg += " row offsets not monotone;";
            if (e & 2) msg += " column out of range;";
            if (e & 4) msg += " non-positive edge weight;";
            if (e & 8) msg += " non-positive vertex weight;";
            if (e & 16) msg += " total weight does not fit in int32;";
            throw InvalidInput(msg);
        }
    }
    return L;
}

struct AttemptResult {
    int64_t cut = 0;
    double t_total = 0, t_coarsen = 0, t_map = 0, t_build = 0, t_init = 0, t_refine = 0;
    std::vector<LevelRecord> recs;
    int64_t largest = 0, smallest = 0, total = 0;
    int32_t opt = 0;
};

static int32_t coarsen_level(const Level& cur, int32_t alg, uint64_t seed, int32_t level, vid_t* map, Arena& a,
                             hipStream_t s) {
    DGraph g = cur.view();
    if (alg == 1) return coarsen_hec(g, cur.uniform, seed, level, map, a, s);
    return coarsen_match(g, cur.uniform, seed, level, alg != 2, map, a, s);
}

// binary_dump.hpp:61-97. Widths of the reference's graph type for the
// entry point used: ordinal, offset and weight bytes (matrix_t 4/4/4,
// big_matrix_t 4/8/8, biggest_matrix_t 8/8/8).
struct DumpWidths {
    int ord, off, val;
};

static void put(FILE* f, int64_t v, int bytes) {
    if (bytes == 4) {
        int32_t x = (int32_t)v;
        std::fwrite(&x, 4, 1, f);
    } else {
        std::fwrite(&v, 8, 1, f);
    }
}

template <class T>
static void put_array(FILE* f, const std::vector<T>& a, int bytes) {
    if ((int)sizeof(T) == bytes) {
        if (!a.empty()) std::fwrite(a.data(), sizeof(T), a.size(), f);
        return;
    }
    const size_t chunk = 1 << 16;
    std::vector<char> buf(chunk * bytes);
    for (size_t i = 0; i < a.size(); i += chunk) {
        const size_t c = std::min(chunk, a.size() - i);
        for (size_t j = 0; j < c; j++) {
            const int64_t v = (int64_t)a[i + j];
            if (bytes == 4) {
                int32_t x = (int32_t)v;
                std::memcpy(buf.data() + 4 * j, &x, 4);
            } else {
                std::memcpy(buf.data() + 8 * j, &v, 8);
            }
        }
        std::fwrite(buf.data(), bytes, c, f);
    }
}

template <class T>
static std::vector<T> to_host(Context& C, const T* d, int64_t count) {
    std::vector<T> h((size_t)count);
    if (count > 0) JET_HIP(hipMemcpyAsync(h.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost, C.stream));
    JET_HIP(hipStreamSynchronize(C.stream));
    return h;
}

// Writes coarse_part.out (the partition of level `last`, int32) and
// coarse_graphs.out (levels 0..last, finest first, each with the map from the
// level before it) in the working directory, as dump_coarse does.
static void dump_hierarchy(Context& C, const std::vector<Level>& levels, size_t last, const part_t* part,
                           const DumpWidths& w) {
    {
        FILE* f = std::fopen("coarse_part.out", "wb");
        if (!f) throw InvalidInput("dump_coarse: cannot write coarse_part.out");
        std::vector<part_t> h = to_host(C, part, levels[last].n);
        put_array(f, h, 4);
        std::fclose(f);
    }
    FILE* f = std::fopen("coarse_graphs.out", "wb");
    if (!f) throw InvalidInput("dump_coarse: cannot write coarse_graphs.out");
    int32_t count = (int32_t)(last + 1);
    std::fwrite(&count, sizeof(int32_t), 1, f);
    for (size_t i = 0; i <= last; i++) {
        const Level& L = levels[i];
        put(f, L.n, w.ord);
        put(f, L.m, w.off);
        put_array(f, to_host(C, L.rowptr, (int64_t)L.n + 1), w.off);
        put_array(f, to_host(C, L.col, L.m), w.ord);
        put_array(f, to_host(C, L.ew, L.m), w.val);
        put_array(f, to_host(C, L.vw, L.n), w.val);
        if (i > 0) {
            const vid_t prev_n = levels[i - 1].n;
            std::vector<vid_t> map;
            if (L.map) {
                map = to_host(C, L.map, prev_n);
            } else {
                map.resize(prev_n);
                for (vid_t v = 0; v < prev_n; v++) map[v] = v;
            }
            put_array(f, map, w.ord);
        }
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    if (!ok) throw InvalidInput("dump_coarse: write to coarse_graphs.out failed");
}

// uncoarsener::multilevel_jet and uncoarsen (uncoarsen.hpp:88-155): refine
// each level from the coarsest (whose partition is `cur`) to the finest and
// project onto the next finer one; the finest result lands in part_out. With
// `dump`, the hierarchy is written before refining the first level whose
// partition is balanced, and the refinement clock restarts there.
static void uncoarsen_levels(Context& C, const jet_config_t& cfg, const std::vector<Level>& levels, part_t* cur,
                             part_t* part_out, AttemptResult& R, const DumpWidths* dump, clk::time_point t_start) {
    hipStream_t s = C.stream;
    const int32_t k = cfg.num_parts;
    const Level& fine = levels.front();
    RefineParams rp;
    rp.k = k;
    rp.imb = cfg.max_imb_ratio;
    rp.tol = cfg.refine_tolerance;
    rp.ultra = cfg.ultra_settings != 0;
    {
        ArenaScope rscope(C.scratch);
        int64_t cap = conn_capacity(fine.n, fine.rowptr, k, C.scratch, s);
        Refiner ref(C.scratch, s, fine.n, cap, rp);
        // A level that shares its graph with the level just refined (the
        // coarsening fixed point) starts from that level's output partition
        // and refine_data. When that level changed neither, this level
        // repeats the same deterministic computation from the same state, so
        // its record is the same and the partition stays as it is.
        bool unchanged = false;
        bool dumped = dump == nullptr;
        LevelRecord last;
        for (size_t L = levels.size(); L-- > 0;) {
            if (!dumped) {
                // uncoarsen.hpp:96-111: the part sizes of `cur` on this level
                // (equal to the refined sizes carried from the coarser level)
                std::vector<int64_t> hs;
                {
                    ArenaScope ds(C.scratch);
                    int64_t* d = C.scratch.alloc<int64_t>(k);
                    part_sizes(levels[L].view(), cur, k, d, s);
                    hs = to_host(C, d, k);
                }
                int64_t mx = 0, tot = 0;
                for (int64_t v : hs) {
                    mx = std::max(mx, v);
                    tot += v;
                }
                const int64_t opt = (tot + k - 1) / k;
                if ((double)mx / (double)opt <= cfg.max_imb_ratio) {
                    dump_hierarchy(C, levels, L, cur, *dump);
                    dumped = true;
                    t_start = clk::now();
                }
            }
            const bool same_graph = L + 1 < levels.size() && levels[L + 1].map == nullptr;
            LevelRecord rec;
            if (same_graph && unchanged) {
                rec = last;
                rec.seconds = 0;
            } else {
                rec = ref.refine(levels[L], cur);
                unchanged = !ref.changed_last();
                last = rec;
            }
            if (cfg.verbose && L + 1 == levels.size())
                std::printf("Initial %lld %.6f %d\n", (long long)(rec.cut2 / 2), 0.0, (int)levels[L].n);
            R.recs.push_back(rec);
            if (L > 0) {
                if (levels[L].map == nullptr && L > 1) continue;  // identity level: same buffer
                part_t* fp = (L == 1) ? part_out : C.persist.alloc<part_t>(levels[L - 1].n);
                if (levels[L].map == nullptr)
                    JET_HIP(hipMemcpyAsync(fp, cur, sizeof(part_t) * levels[L].n, hipMemcpyDeviceToDevice, s));
                else
                    project(levels[L - 1].n, levels[L].map, cur, fp, s);
                cur = fp;
            }
        }
        R.cut = ref.cut2() / 2;
        R.total = ref.total_size();
        R.opt = ref.opt();
        std::vector<int64_t> hs(k);
        JET_HIP(hipMemcpyAsync(hs.data(), ref.best_sizes_dev(), sizeof(int64_t) * k, hipMemcpyDeviceToHost, s));
        JET_HIP(hipStreamSynchronize(s));
        R.largest = 0;
        R.smallest = R.total;
        for (int64_t v : hs) {
            R.largest = std::max(R.largest, v);
            R.smallest = std::min(R.smallest, v);
        }
    }
    R.t_refine = secs(t_start, clk::now());
}

// The initial partition of a large coarsest graph `cs`, multilevel k-way
// (initpart.h): the host coarsens it further, partitions the smallest level
// by recursive bisection, and each level of that hierarchy is refined on the
// device with Jet refinement (its own refine_data) before its partition is
// projected to the next finer one; the partition of `cs` itself is left to
// the uncoarsening that follows. Levels live in `persist`, the refinement
// state in `scratch`.
static void initial_partition_multilevel(Context& C, const jet_config_t& cfg, const Level& cs,
                                         const std::vector<int64_t>& h_rp, const std::vector<int32_t>& h_col,
                                         const std::vector<int32_t>& h_ew, const std::vector<int32_t>& h_vw,
                                         uint64_t seed, part_t* cs_part) {
    hipStream_t s = C.stream;
    const int32_t k = cfg.num_parts;
    std::vector<jet_init::HostLevel> H =
        jet_init::init_hierarchy(cs.n, h_rp.data(), h_col.data(), h_ew.data(), h_vw.data(), k, seed);
    std::vector<int32_t> hp;
    if (H.empty()) {  // no coarser level: the recursive bisection alone
        hp.resize(cs.n);
        if (jet_host_initpart(cs.n, h_rp.data(), h_col.data(), h_ew.data(), h_vw.data(), k, cfg.max_imb_ratio, seed,
                              hp.data()) != 0)
            throw InitFailure("initial partitioner could not partition the coarsest graph");
        JET_HIP(hipMemcpyAsync(cs_part, hp.data(), sizeof(int32_t) * cs.n, hipMemcpyHostToDevice, s));
        return;
    }
    const jet_init::HostLevel& last = H.back();
    hp.resize(last.n);
    if (jet_host_initpart(last.n, last.xadj.data(), last.adj.data(), last.adjw.data(), last.vw.data(), k,
                          cfg.max_imb_ratio, seed, hp.data()) != 0)
        throw InitFailure("initial partitioner could not partition the coarsest graph");
    // the hierarchy's levels and maps on the device
    std::vector<Level> D(H.size());
    for (size_t i = 0; i < H.size(); i++) {
        const jet_init::HostLevel& h = H[i];
        Level& d = D[i];
        d.n = h.n;
        d.m = h.xadj[h.n];
        d.rowptr = C.persist.alloc<eoff_t>((int64_t)h.n + 1);
        d.col = C.persist.alloc<vid_t>(std::max<int64_t>(d.m, 1));
        d.ew = C.persist.alloc<wgt_t>(std::max<int64_t>(d.m, 1));
        d.vw = C.persist.alloc<wgt_t>(std::max<int64_t>(h.n, 1));
        const int64_t nf = i == 0 ? cs.n : H[i - 1].n;
        d.map = C.persist.alloc<vid_t>(std::max<int64_t>(nf, 1));
        d.uniform = false;
        JET_HIP(hipMemcpyAsync(d.rowptr, h.xadj.data(), sizeof(int64_t) * (h.n + 1), hipMemcpyHostToDevice, s));
        if (d.m > 0) {
            JET_HIP(hipMemcpyAsync(d.col, h.adj.data(), sizeof(int32_t) * d.m, hipMemcpyHostToDevice, s));
            JET_HIP(hipMemcpyAsync(d.ew, h.adjw.data(), sizeof(int32_t) * d.m, hipMemcpyHostToDevice, s));
        }
        JET_HIP(hipMemcpyAsync(d.vw, h.vw.data(), sizeof(int32_t) * h.n, hipMemcpyHostToDevice, s));
        JET_HIP(hipMemcpyAsync(d.map, h.map.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, s));
    }
    RefineParams rp;
    rp.k = k;
    rp.imb = cfg.max_imb_ratio;
    rp.tol = cfg.refine_tolerance;
    rp.ultra = cfg.ultra_settings != 0;
    ArenaScope rscope(C.scratch);
    int64_t cap = conn_capacity(D[0].n, D[0].rowptr, k, C.scratch, s);
    Refiner ref(C.scratch, s, D[0].n, cap, rp);
    part_t* cur = C.persist.alloc<part_t>(last.n);
    JET_HIP(hipMemcpyAsync(cur, hp.data(), sizeof(int32_t) * last.n, hipMemcpyHostToDevice, s));
    for (size_t L = D.size(); L-- > 0;) {
        ref.refine(D[L], cur);
        part_t* fp = L == 0 ? cs_part : C.persist.alloc<part_t>(H[L - 1].n);
        project(L == 0 ? cs.n : H[L - 1].n, D[L].map, cur, fp, s);
        cur = fp;
    }
}

// partitioner.hpp:61-125. part_out: device buffer of fine.n entries.
static AttemptResult partition_once(Context& C, const jet_config_t& cfg, const Level& fine, uint64_t seed,
                                    part_t* part_out, const DumpWidths* dump) {
    hipStream_t s = C.stream;
    AttemptResult R;
    auto t0 = clk::now();
    const int32_t k = cfg.num_parts;
    int32_t alg = cfg.coarsening_alg;
    if (alg != 1 && alg != 2) alg = 0;  // default heuristic: MtMetis (2-hop matching)
    int32_t cutoff = k * 8;
    if (cutoff > 1024) {
        cutoff = k * 2;
        cutoff = std::max(1024, cutoff);
    }
    const int32_t min_allowed = cutoff / 4;
    Arena::Mark level_mark = C.persist.mark();
    std::vector<Level> levels{fine};
    bool fixed_point = false;
    while (levels.back().n > cutoff) {
        Level& cur = levels.back();
        if (fixed_point) {
            // identity level: same arrays, map == nullptr means map[i] = i
            Level same = cur;
            same.map = nullptr;
            levels.push_back(same);
            if (levels.size() > 200) break;
            continue;
        }
        auto tm = clk::now();
        vid_t* map = C.persist.alloc<vid_t>(cur.n);
        int32_t nc = coarsen_level(cur, alg, seed, (int32_t)levels.size() - 1, map, C.scratch, s);
        JET_HIP(hipStreamSynchronize(s));
        auto tb = clk::now();
        R.t_map += secs(tm, tb);
        if (nc < min_allowed) break;
        // Weighted HEC draws no random numbers and ignores the level index, and
        // map ids follow root order, so nc == n means map[i] = i. Contracting a
        // contracted level (rows sorted, no self-loops or duplicates) through
        // the identity returns the same graph, so every later level up to the
        // level limit repeats it: they share its arrays instead of being
        // rebuilt (contract.hpp:508-527 would rebuild the same graph each time).
        if (alg == 1 && levels.size() >= 2 && nc == cur.n) {
            fixed_point = true;
            Level same = cur;
            same.map = nullptr;
            R.t_build += secs(tb, clk::now());
            levels.push_back(same);
            if (levels.size() > 200) break;
            continue;
        }
        Level next;
        contract(cur.view(), map, nc, C.persist, C.scratch, s, next);
        next.map = map;
        next.uniform = false;
        JET_HIP(hipStreamSynchronize(s));
        R.t_build += secs(tb, clk::now());
        levels.push_back(next);
        if (levels.size() > 200) break;
    }
    auto t1 = clk::now();
    R.t_coarsen = secs(t0, t1);

    // initial partition of the coarsest graph on the host
    const Level& cs = levels.back();
    std::vector<int64_t> h_rp(cs.n + 1);
    std::vector<int32_t> h_col(cs.m), h_ew(cs.m), h_vw(cs.n), h_part(cs.n);
    JET_HIP(hipMemcpyAsync(h_rp.data(), cs.rowptr, sizeof(int64_t) * (cs.n + 1), hipMemcpyDeviceToHost, s));
    if (cs.m > 0) {
        JET_HIP(hipMemcpyAsync(h_col.data(), cs.col, sizeof(int32_t) * cs.m, hipMemcpyDeviceToHost, s));
        JET_HIP(hipMemcpyAsync(h_ew.data(), cs.ew, sizeof(int32_t) * cs.m, hipMemcpyDeviceToHost, s));
    }
    JET_HIP(hipMemcpyAsync(h_vw.data(), cs.vw, sizeof(int32_t) * cs.n, hipMemcpyDeviceToHost, s));
    JET_HIP(hipStreamSynchronize(s));
    part_t* cur = (levels.size() == 1) ? part_out : C.persist.alloc<part_t>(cs.n);
    if (jet_init::wants_hierarchy(cs.n, cs.m, k)) {
        initial_partition_multilevel(C, cfg, cs, h_rp, h_col, h_ew, h_vw, seed, cur);
    } else {
        if (jet_host_initpart(cs.n, h_rp.data(), h_col.data(), h_ew.data(), h_vw.data(), k, cfg.max_imb_ratio, seed,
                              h_part.data()) != 0)
            throw InitFailure("initial partitioner could not partition the coarsest graph");
        JET_HIP(hipMemcpyAsync(cur, h_part.data(), sizeof(int32_t) * cs.n, hipMemcpyHostToDevice, s));
    }
    JET_HIP(hipStreamSynchronize(s));
    auto t2 = clk::now();
    R.t_init = secs(t1, t2);

    // uncoarsening (uncoarsen.hpp:88-125)
    uncoarsen_levels(C, cfg, levels, cur, part_out, R, dump, t2);
    R.t_total = secs(t0, clk::now());
    C.persist.release(level_mark);
    return R;
}

static void fill_stats(const AttemptResult& R, int32_t attempts, int32_t k, jet_stats_t* st) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->total_s = R.t_total;
    st->coarsen_s = R.t_coarsen;
    st->map_s = R.t_map;
    st->build_s = R.t_build;
    st->init_partition_s = R.t_init;
    st->refine_s = R.t_refine;
    st->attempts = attempts;
    st->num_levels = (int32_t)std::min<size_t>(R.recs.size(), JET_MAX_LEVELS);
    for (int32_t i = 0; i < st->num_levels; i++) {
        const LevelRecord& r = R.recs[i];
        jet_level_stats_t& o = st->levels[i];
        o.edge_cut = r.cut2 / 2;
        o.imb = R.opt > 0 ? (double)r.total_imb / (double)R.opt : 0.0;
        o.num_edges = r.m;
        o.num_vertices = r.n;
        o.total_ref_time = r.seconds;
        o.total_iterations = r.iterations;
        o.lp_iterations = r.lp_iterations;
    }
    double opt = (double)((R.total + k - 1) / k);
    st->finest_imb_ratio = opt > 0 ? (double)R.largest / opt : 0.0;
    st->finest_edge_cut = R.cut;
    st->largest_part = R.largest;
    st->smallest_part = R.smallest;
}

static void check_config(const jet_config_t* cfg) {
    if (!cfg) throw InvalidInput("config is NULL");
    if (cfg->num_parts < 1 || cfg->num_parts > JET_MAX_PARTS) throw InvalidInput("num_parts must be in [1, 8192]");
    if (!(cfg->max_imb_ratio >= 1.0)) throw InvalidInput("max_imb_ratio must be >= 1");
}

// Attempt loop of app/driver.cpp:110-130: best cut by strict <, seeds seed + i.
static void partition_entry(const jet_config_t* cfg, int64_t n, const int64_t* rp64, const int32_t* rp32,
                            const int32_t* col, const int32_t* ew32, const int64_t* ew64, const int32_t* vw32,
                            const int64_t* vw64, int32_t uniform_ew, int32_t space, uint64_t seed, int32_t* part_out,
                            int64_t* cut_out, jet_stats_t* stats, const int64_t* col64 = nullptr,
                            Context* ctx = nullptr) {
    check_config(cfg);
    if (!part_out || !cut_out) throw InvalidInput("part_out and cut_out are required");
    if (cfg->dump_coarse && ctx) throw InvalidInput("dump_coarse is not supported for jobs of a jet_ctx");
    // the reference's graph type for this entry point sets the dump's widths
    const DumpWidths widths{col64 ? 8 : 4, rp32 ? 4 : 8, rp32 ? 4 : 8};
    const DumpWidths* dump = cfg->dump_coarse ? &widths : nullptr;
    Context& C = ctx ? *ctx : context();
    CallScope call(C);
    Level fine = stage_graph(C, n, rp64, rp32, col, ew32, ew64, vw32, vw64, space, uniform_ew != 0, true, col64);
    int32_t attempts = std::max(cfg->num_iter, 1);
    part_t* best = (space == JET_PTR_DEVICE && attempts == 1) ? part_out : C.persist.alloc<part_t>(n);
    part_t* trial = attempts > 1 ? C.persist.alloc<part_t>(n) : best;
    AttemptResult best_r;
    bool have = false;
    for (int32_t i = 0; i < attempts; i++) {
        AttemptResult r = partition_once(C, *cfg, fine, seed + (uint64_t)i, trial, dump);
        if (!have || r.cut < best_r.cut) {
            if (trial != best) copy_i32(best, trial, n, C.stream);
            best_r = r;
            have = true;
        }
    }
    from_device(C, part_out, best, n, space);
    *cut_out = best_r.cut;
    fill_stats(best_r, attempts, cfg->num_parts, stats);
}

}  // namespace jet

using namespace jet;

namespace jet {
static int guard_impl(const std::function<void()>& f) {
    try {
        f();
        g_err.clear();
        return JET_OK;
    } catch (const InvalidInput& e) {
        g_err = e.what();
        return JET_EINVAL;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return JET_EINVAL;
    } catch (const InitFailure& e) {
        g_err = e.what();
        return JET_EINIT;
    } catch (const std::bad_alloc&) {
        g_err = "device memory allocation failed";
        return JET_ENOMEM;
    } catch (const DeviceError& e) {
        g_err = e.what();
        return JET_EDEVICE;
    } catch (const std::exception& e) {
        g_err = e.what();
        return JET_EDEVICE;
    }
}
int partition_in_context(Context& C, const jet_config_t* cfg, int64_t n, const int64_t* row_offsets,
                         const int32_t* col, const int32_t* ew, const int32_t* vw, int32_t uniform_ew,
                         int32_t ptr_space, uint64_t seed, int32_t* part_out, int64_t* cut_out, jet_stats_t* stats) {
    return guard_impl([&] {
        partition_entry(cfg, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, uniform_ew, ptr_space, seed,
                        part_out, cut_out, stats, nullptr, &C);
    });
}

void set_last_error(const std::string& msg) { g_err = msg; }

}  // namespace jet

extern "C" {

void jet_config_default(jet_config_t* c) {
    if (!c) return;
    c->coarsening_alg = 0;
    c->num_iter = 1;
    c->max_imb_ratio = 1.03;
    c->num_parts = 2;
    c->refine_tolerance = 0.999;
    c->dump_coarse = 0;
    c->verbose = 0;
    c->ultra_settings = 0;
}

const char* jet_last_error(void) { return g_err.c_str(); }

int jet_device_count(void) {
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return cnt;
}

int jet_partition_csr(const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int32_t* col,
                      const int32_t* ew, const int32_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
                      int32_t* part_out, int64_t* cut_out, jet_stats_t* stats) {
    return guard_impl([&] {
        partition_entry(cfg, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, uniform_ew, ptr_space, seed,
                        part_out, cut_out, stats);
    });
}

int jet_partition_csr32(const jet_config_t* cfg, int64_t n, const int32_t* row_offsets, const int32_t* col,
                        const int32_t* ew, const int32_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
                        int32_t* part_out, int64_t* cut_out, jet_stats_t* stats) {
    return guard_impl([&] {
        partition_entry(cfg, n, nullptr, row_offsets, col, ew, nullptr, vw, nullptr, uniform_ew, ptr_space, seed,
                        part_out, cut_out, stats);
    });
}

int jet_partition_csr_big(const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int32_t* col,
                          const int64_t* ew, const int64_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
                          int32_t* part_out, int64_t* cut_out, jet_stats_t* stats) {
    return guard_impl([&] {
        partition_entry(cfg, n, row_offsets, nullptr, col, nullptr, ew, nullptr, vw, uniform_ew, ptr_space, seed,
                        part_out, cut_out, stats);
    });
}

int jet_partition_csr_biggest(const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int64_t* col,
                              const int64_t* ew, const int64_t* vw, int32_t uniform_ew, int32_t ptr_space,
                              uint64_t seed, int32_t* part_out, int64_t* cut_out, jet_stats_t* stats) {
    return guard_impl([&] {
        partition_entry(cfg, n, row_offsets, nullptr, nullptr, nullptr, ew, nullptr, vw, uniform_ew, ptr_space, seed,
                        part_out, cut_out, stats, col);
    });
}

int jet_coarsen_map(int32_t alg, int32_t level, int32_t uniform, uint64_t seed, int64_t n, const int64_t* row_offsets,
                    const int32_t* col, const int32_t* ew, const int32_t* vw, int32_t ptr_space, int32_t* map_out,
                    int32_t* nc_out) {
    return guard_impl([&] {
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, ptr_space, uniform != 0);
        vid_t* map = ptr_space == JET_PTR_DEVICE ? map_out : C.persist.alloc<vid_t>(n);
        int32_t a = (alg == 1 || alg == 2) ? alg : 0;
        *nc_out = coarsen_level(L, a, seed, level, map, C.scratch, C.stream);
        from_device(C, map_out, map, n, ptr_space);
    });
}

int jet_hec_heavy_neighbours(int32_t level, int32_t uniform, uint64_t seed, int64_t n, const int64_t* row_offsets,
                             const int32_t* col, const int32_t* ew, const int32_t* vw, int32_t ptr_space,
                             int32_t* hn_out) {
    return guard_impl([&] {
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, ptr_space, uniform != 0);
        vid_t* hn = ptr_space == JET_PTR_DEVICE ? hn_out : C.persist.alloc<vid_t>(n);
        hec_heavy_neighbours(L.view(), uniform != 0, seed, level, hn, C.scratch, C.stream);
        from_device(C, hn_out, hn, n, ptr_space);
    });
}

int jet_map_construct(int64_t n, const int32_t* hn, int32_t ptr_space, int32_t* map_out, int32_t* nc_out) {
    return guard_impl([&] {
        check_space(ptr_space);
        if (n < 1 || n > ORD_MAX - 2) throw InvalidInput("n out of range");
        Context& C = context();
        CallScope call(C);
        const vid_t* dh = to_device(C, hn, n, ptr_space);
        // the claim rounds index stamp[hn[u]] and claim[hn[u]]
        if (!parts_in_range(n, dh, (int32_t)n, C.scratch, C.stream)) throw InvalidInput("hn entries must be in [0, n)");
        vid_t* map = ptr_space == JET_PTR_DEVICE ? map_out : C.persist.alloc<vid_t>(n);
        *nc_out = map_construct((vid_t)n, dh, map, C.scratch, C.stream);
        from_device(C, map_out, map, n, ptr_space);
    });
}

int jet_contract(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* vw,
                 const int32_t* map, int32_t nc, int32_t ptr_space, int64_t* out_row_offsets, int32_t* out_col,
                 int32_t* out_ew, int32_t* out_vw, int64_t* out_m) {
    return guard_impl([&] {
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, ptr_space, false);
        if (nc < 1 || nc > n) throw InvalidInput("nc must be in [1, n]");
        const vid_t* dmap = to_device(C, map, n, ptr_space);
        // the counting pass adds into per-coarse-vertex arrays at map[i]
        if (!parts_in_range(n, dmap, nc, C.scratch, C.stream)) throw InvalidInput("map entries must be in [0, nc)");
        Level out;
        contract(L.view(), dmap, nc, C.persist, C.scratch, C.stream, out);
        from_device(C, out_row_offsets, out.rowptr, (int64_t)nc + 1, ptr_space);
        from_device(C, out_col, out.col, out.m, ptr_space);
        from_device(C, out_ew, out.ew, out.m, ptr_space);
        from_device(C, out_vw, out.vw, nc, ptr_space);
        *out_m = out.m;
    });
}

int jet_conn_tables(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                    const int32_t* part, int32_t k, int32_t ptr_space, int64_t* out_off, int32_t* out_part,
                    int32_t* out_val, int32_t* out_size) {
    return guard_impl([&] {
        if (ptr_space != JET_PTR_HOST) throw InvalidInput("jet_conn_tables takes host pointers");
        check_k(k);
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, nullptr, nullptr, ptr_space, false);
        const part_t* dp = to_device(C, part, n, ptr_space);
        check_parts(C, n, dp, k);
        RefineParams rp;
        rp.k = k;
        int64_t cap = conn_capacity(L.n, L.rowptr, k, C.scratch, C.stream);
        Refiner ref(C.scratch, C.stream, L.n, cap, rp);
        std::vector<int64_t> off;
        std::vector<int32_t> sp, sv, sz;
        ref.conn_tables(L, dp, off, sp, sv, sz);
        std::copy(off.begin(), off.end(), out_off);
        std::copy(sp.begin(), sp.end(), out_part);
        std::copy(sv.begin(), sv.end(), out_val);
        std::copy(sz.begin(), sz.end(), out_size);
    });
}

int jet_refine_step(int32_t kind, int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                    const int32_t* vw, const int32_t* part, int32_t k, double imb, double filter_ratio,
                    int32_t ptr_space, int32_t* moves_out, int32_t* dest_out, int32_t* nmoves_out) {
    return guard_impl([&] {
        if (ptr_space != JET_PTR_HOST) throw InvalidInput("jet_refine_step takes host pointers");
        if (kind < 0 || kind > 2) throw InvalidInput("kind must be 0, 1 or 2");
        check_k(k);
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, ptr_space, false);
        const part_t* dp = to_device(C, part, n, ptr_space);
        check_parts(C, n, dp, k);
        RefineParams rp;
        rp.k = k;
        rp.imb = imb;
        int64_t cap = conn_capacity(L.n, L.rowptr, k, C.scratch, C.stream);
        Refiner ref(C.scratch, C.stream, L.n, cap, rp);
        std::vector<int32_t> mv, de;
        ref.single_step(L, dp, kind, filter_ratio, mv, de);
        std::copy(mv.begin(), mv.end(), moves_out);
        std::copy(de.begin(), de.end(), dest_out);
        *nmoves_out = (int32_t)mv.size();
    });
}

int jet_refine_level(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* vw,
                     int32_t k, double imb, double tol, int32_t ultra, int32_t uniform, int32_t ptr_space,
                     int32_t* part_io, int64_t* rec_out) {
    return jet_refine_level_trace(n, row_offsets, col, ew, vw, k, imb, tol, ultra, uniform, ptr_space, part_io, rec_out,
                                  nullptr, 0, nullptr);
}

int jet_refine_level_trace(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                           const int32_t* vw, int32_t k, double imb, double tol, int32_t ultra, int32_t uniform,
                           int32_t ptr_space, int32_t* part_io, int64_t* rec_out, int64_t* trace_out,
                           int32_t max_iters, int32_t* iters_out) {
    return guard_impl([&] {
        check_k(k);
        if (iters_out && (max_iters < 0 || (max_iters > 0 && !trace_out)))
            throw InvalidInput("trace_out must hold max_iters >= 0 records");
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, ptr_space, uniform != 0);
        part_t* dp = C.persist.alloc<part_t>(n);
        if (ptr_space == JET_PTR_DEVICE)
            JET_HIP(hipMemcpyAsync(dp, part_io, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, C.stream));
        else JET_HIP(hipMemcpyAsync(dp, part_io, sizeof(int32_t) * n, hipMemcpyHostToDevice, C.stream));
        check_parts(C, n, dp, k);
        RefineParams rp;
        rp.k = k;
        rp.imb = imb;
        rp.tol = tol;
        rp.ultra = ultra != 0;
        int64_t cap = conn_capacity(L.n, L.rowptr, k, C.scratch, C.stream);
        Refiner ref(C.scratch, C.stream, L.n, cap, rp);
        std::vector<std::array<int64_t, 6>> tr;
        if (iters_out) ref.trace = &tr;
        LevelRecord rec = ref.refine(L, dp);
        from_device(C, part_io, dp, n, ptr_space);
        if (iters_out) {
            *iters_out = (int32_t)tr.size();
            for (size_t t = 0; t < tr.size() && (int64_t)t < max_iters; t++)
                for (int c = 0; c < 6; c++) trace_out[6 * t + c] = tr[t][c];
        }
        rec_out[0] = rec.cut2;
        rec_out[1] = rec.total_imb;
        rec_out[2] = rec.n;
        rec_out[3] = rec.m;
        rec_out[4] = rec.iterations;
        rec_out[5] = rec.lp_iterations;
        rec_out[6] = ref.opt();
        rec_out[7] = ref.size_max();
    });
}

int jet_project(int64_t n_fine, const int32_t* map, const int32_t* coarse, int32_t ptr_space, int32_t* fine_out) {
    return guard_impl([&] {
        check_space(ptr_space);
        if (n_fine < 0 || n_fine > ORD_MAX - 2) throw InvalidInput("n out of range");
        Context& C = context();
        CallScope call(C);
        const vid_t* dmap = to_device(C, map, n_fine, ptr_space);
        int32_t nc = 0;
        if (ptr_space == JET_PTR_HOST) {
            for (int64_t i = 0; i < n_fine; i++) {
                if (map[i] < 0) throw InvalidInput("map entries must be >= 0");
                nc = std::max(nc, map[i] + 1);
            }
        } else {
            throw InvalidInput("jet_project with device pointers needs the coarse size; use host pointers");
        }
        const part_t* dc = to_device(C, coarse, nc, ptr_space);
        part_t* df = C.persist.alloc<part_t>(n_fine);
        project((vid_t)n_fine, dmap, dc, df, C.stream);
        from_device(C, fine_out, df, n_fine, ptr_space);
    });
}

int jet_edge_cut(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* part,
                 int32_t ptr_space, int64_t* cut2_out) {
    return guard_impl([&] {
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, nullptr, nullptr, ptr_space, false);
        const part_t* dp = to_device(C, part, n, ptr_space);
        *cut2_out = total_cut2(L.view(), dp, C.scratch, C.stream);
    });
}

int jet_part_sizes(int64_t n, const int32_t* vw, const int32_t* part, int32_t k, int32_t ptr_space,
                   int64_t* sizes_out) {
    return guard_impl([&] {
        check_space(ptr_space);
        if (n < 1 || k < 1 || k > JET_MAX_PARTS) throw InvalidInput("bad n or k");
        Context& C = context();
        CallScope call(C);
        DGraph g;
        g.n = (vid_t)n;
        wgt_t* dvw;
        if (vw) {
            dvw = to_device(C, vw, n, ptr_space);
        } else {
            dvw = C.persist.alloc<wgt_t>(n);
            fill_i32(dvw, 1, n, C.stream);
        }
        g.vw = dvw;
        const part_t* dp = to_device(C, part, n, ptr_space);
        check_parts(C, n, dp, k);
        int64_t* ds = ptr_space == JET_PTR_DEVICE ? sizes_out : C.persist.alloc<int64_t>(k);
        part_sizes(g, dp, k, ds, C.stream);
        from_device(C, sizes_out, ds, k, ptr_space);
    });
}

int jet_uncoarsen_hierarchy(const jet_config_t* cfg, const jet_hierarchy_t* h, const int32_t* coarse_part,
                            int32_t* part_out, int64_t* cut_out, jet_stats_t* stats) {
    return guard_impl([&] {
        check_config(cfg);
        if (cfg->dump_coarse) throw InvalidInput("dump_coarse must be 0 when replaying a hierarchy");
        if (!h || h->levels.empty() || !coarse_part || !part_out || !cut_out)
            throw InvalidInput("hierarchy, coarse_part, part_out and cut_out are required");
        Context& C = context();
        CallScope call(C);
        auto t0 = clk::now();
        std::vector<Level> levels;
        for (size_t l = 0; l < h->levels.size(); l++) {
            const jet_hierarchy::Level& H = h->levels[l];
            if (l > 0 && (int64_t)H.map.size() != h->levels[l - 1].n) throw InvalidInput("map length mismatch");
            Level L = stage_graph(C, H.n, H.rowptr.data(), nullptr, H.col.data(), H.ew.data(), nullptr, H.vw.data(),
                                  nullptr, JET_PTR_HOST, false);
            if (l > 0) L.map = to_device(C, H.map.data(), (int64_t)H.map.size(), JET_PTR_HOST);
            levels.push_back(L);
        }
        const Level& cs = levels.back();
        part_t* cur = C.persist.alloc<part_t>(cs.n);
        JET_HIP(hipMemcpyAsync(cur, coarse_part, sizeof(int32_t) * cs.n, hipMemcpyHostToDevice, C.stream));
        check_parts(C, cs.n, cur, cfg->num_parts);
        const vid_t fine_n = levels.front().n;
        part_t* fine_part = levels.size() == 1 ? cur : C.persist.alloc<part_t>(fine_n);
        AttemptResult R;
        auto t1 = clk::now();
        R.t_init = secs(t0, t1);
        uncoarsen_levels(C, *cfg, levels, cur, fine_part, R, nullptr, t1);
        R.t_total = secs(t0, clk::now());
        from_device(C, part_out, fine_part, fine_n, JET_PTR_HOST);
        *cut_out = R.cut;
        fill_stats(R, 1, cfg->num_parts, stats);
    });
}

int jet_partition_quality(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                          const int32_t* vw, const int32_t* part, int32_t k, int32_t ptr_space, jet_quality_t* out) {
    return guard_impl([&] {
        if (!out) throw InvalidInput("out is NULL");
        check_k(k);
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, vw, nullptr, ptr_space, false);
        const part_t* dp = to_device(C, part, n, ptr_space);
        check_parts(C, n, dp, k);
        const int64_t cut2 = total_cut2(L.view(), dp, C.scratch, C.stream);
        int64_t* ds = C.persist.alloc<int64_t>(k);
        part_sizes(L.view(), dp, k, ds, C.stream);
        std::vector<int64_t> hs(k);
        JET_HIP(hipMemcpyAsync(hs.data(), ds, sizeof(int64_t) * k, hipMemcpyDeviceToHost, C.stream));
        RefineParams rp;
        rp.k = k;
        int64_t cap = conn_capacity(L.n, L.rowptr, k, C.scratch, C.stream);
        Refiner ref(C.scratch, C.stream, L.n, cap, rp);
        int64_t comm = 0;
        std::vector<int64_t> cpp;
        ref.quality(L, dp, comm, cpp);
        int64_t total = 0, largest = 0, smallest = INT64_MAX;
        for (int64_t v : hs) {
            total += v;
            largest = std::max(largest, v);
            smallest = std::min(smallest, v);
        }
        const double opt = (double)((total + k - 1) / k);
        out->edge_cut = cut2 / 2;
        out->comm_size = comm;
        out->max_part_cut = cpp.empty() ? 0 : *std::max_element(cpp.begin(), cpp.end());
        out->largest_part = largest;
        out->smallest_part = smallest;
        out->max_imb = opt > 0 ? (double)largest / opt : 0.0;
        out->min_imb = opt > 0 ? (double)smallest / opt : 0.0;
    });
}

int jet_cut_heatmap(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* part,
                    int32_t k, int32_t ptr_space, int64_t* heat_out) {
    return guard_impl([&] {
        if (!heat_out) throw InvalidInput("heat_out is NULL");
        check_k(k);
        Context& C = context();
        CallScope call(C);
        Level L = stage_graph(C, n, row_offsets, nullptr, col, ew, nullptr, nullptr, nullptr, ptr_space, false);
        const part_t* dp = to_device(C, part, n, ptr_space);
        check_parts(C, n, dp, k);
        const int64_t kk = (int64_t)k * k;
        int64_t* dh = ptr_space == JET_PTR_DEVICE ? heat_out : C.persist.alloc<int64_t>(kk);
        RefineParams rp;
        rp.k = k;
        int64_t cap = conn_capacity(L.n, L.rowptr, k, C.scratch, C.stream);
        Refiner ref(C.scratch, C.stream, L.n, cap, rp);
        ref.heatmap(L, dp, dh);
        from_device(C, heat_out, dh, kk, ptr_space);
    });
}

int jet_gain_bucket_batch(int64_t count, const int64_t* gains, const int64_t* vwgts, int32_t* out) {
    return guard_impl([&] {
        if (count < 0) throw InvalidInput("count < 0");
        Context& C = context();
        CallScope call(C);
        const int64_t* dg = to_device(C, gains, count, JET_PTR_HOST);
        const int64_t* dw = to_device(C, vwgts, count, JET_PTR_HOST);
        int32_t* dout = C.persist.alloc<int32_t>(count);
        gain_bucket_batch(count, dg, dw, dout, C.stream);
        from_device(C, out, dout, count, JET_PTR_HOST);
    });
}

int jet_prof_enable(const char* family) {
    return guard_impl([&] {
        if (!family) throw InvalidInput("family is NULL");
        context();
        if (!prof_enable(family)) throw InvalidInput(std::string("unknown kernel family: ") + family);
    });
}

int jet_prof_read(const char* family, int64_t* units, double* total_ms, double* bytes) {
    return guard_impl([&] {
        if (!family || !units || !total_ms || !bytes) throw InvalidInput("family and outputs are required");
        int fam = -1;
        for (int f = 0; f < PF_FAMILIES; f++)
            if (std::strcmp(family, kProfNames[f]) == 0) fam = f;
        if (fam < 0) throw InvalidInput(std::string("unknown kernel family: ") + family);
        int64_t l = 0;
        double ms = 0, b = 0;
        if (!prof_read(fam, l, ms, b)) throw InvalidInput(std::string("family not enabled: ") + family);
        *units = l;
        *total_ms = ms;
        *bytes = b;
    });
}

}  // extern "C"
