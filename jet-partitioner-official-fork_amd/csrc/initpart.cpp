// Host initial partitioner for the coarsest graph.
//
// The reference calls METIS_PartGraphKway on the coarsest level
// (initial_partition.hpp:92-114) after narrowing every array to `int`. METIS
// is not available to this build, so this file provides a deterministic
// multi-trial recursive bisection in its place: greedy graph growing from
// several seeds, then balance-constrained two-way Fiduccia-Mattheyses passes,
// recursing on the induced halves with part counts split floor/ceil. The
// per-bisection tolerance is imb^(1/ceil(log2 k)), so the product of the
// tolerances along any root-to-leaf path stays within `imb` (METIS's ubvec).
//
// It runs on the host in both the GPU path and the CPU oracle so that the two
// start refinement from the same coarsest partition. Independent work (the two
// halves of a bisection, the trials of one bisection) runs on up to
// JET_HOST_THREADS_MAX host threads; every thread writes its own output and the
// best trial is picked in trial order, so the result does not depend on the
// thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <functional>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>

#include <sched.h>
#include <sys/mman.h>

#include "initpart.h"

namespace {

// Blocks of 2 MiB or more for BigVec: mapped with transparent huge pages
// requested, and kept after release for reuse by later arrays of a similar
// size (best fit within 2x), since first-touch page faults, not arithmetic,
// dominated bisections of graphs with tens of millions of edges. Emptied at
// the end of every jet_host_initpart call.
class BlockPool {
  public:
    static constexpr size_t HUGE_BYTES = size_t(2) << 20;
    void* take(size_t bytes, size_t& got) {
        bytes = (bytes + HUGE_BYTES - 1) / HUGE_BYTES * HUGE_BYTES;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= 2 * bytes) {
                got = it->first;
                void* p = it->second;
                free_.erase(it);
                return p;
            }
        }
        void* m = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (m == MAP_FAILED) throw std::bad_alloc();
        (void)madvise(m, bytes, MADV_HUGEPAGE);
        got = bytes;
        return m;
    }
    void give(void* p, size_t bytes) {
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(bytes, p);
    }
    void trim() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& kv : free_) munmap(kv.second, kv.first);
        free_.clear();
    }

  private:
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
};

BlockPool& block_pool() {
    static BlockPool* p = new BlockPool();
    return *p;
}

// Owned array for the large per-graph arrays: no zero fill on resize; blocks
// of 2 MiB or more come from the block pool.
template <class T>
class BigVec {
  public:
    BigVec() = default;
    BigVec(const BigVec&) = delete;
    BigVec& operator=(const BigVec&) = delete;
    BigVec(BigVec&& o) noexcept { swap(o); }
    BigVec& operator=(BigVec&& o) noexcept {
        BigVec t(std::move(o));
        swap(t);
        return *this;
    }
    ~BigVec() { release(); }

    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    T* data() { return p_; }
    const T* data() const { return p_; }
    T& operator[](size_t i) { return p_[i]; }
    const T& operator[](size_t i) const { return p_[i]; }
    T* begin() { return p_; }
    T* end() { return p_ + n_; }
    const T* begin() const { return p_; }
    const T* end() const { return p_ + n_; }

    // Contents beyond the old size are unspecified.
    void resize(size_t n) {
        if (n > cap_) grow(n);
        n_ = n;
    }
    void assign(size_t n, T v) {
        resize(n);
        std::fill(p_, p_ + n, v);
    }
    void assign(const T* b, const T* e) {
        resize((size_t)(e - b));
        std::memcpy(p_, b, sizeof(T) * n_);
    }

  private:
    static constexpr size_t HUGE_BYTES = BlockPool::HUGE_BYTES;
    void swap(BigVec& o) {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        std::swap(cap_, o.cap_);
        std::swap(mapped_, o.mapped_);
    }
    void grow(size_t cap) {
        size_t bytes = sizeof(T) * cap;
        T* q = nullptr;
        bool mapped = false;
        if (bytes >= HUGE_BYTES) {
            size_t got = 0;
            q = static_cast<T*>(block_pool().take(bytes, got));
            mapped = true;
            cap = got / sizeof(T);
        } else {
            q = static_cast<T*>(std::malloc(bytes > 0 ? bytes : 1));
            if (!q) throw std::bad_alloc();
        }
        if (n_ > 0) std::memcpy(q, p_, sizeof(T) * n_);
        release();
        p_ = q;
        cap_ = cap;
        mapped_ = mapped;
    }
    void release() {
        if (!p_) return;
        if (mapped_) block_pool().give(p_, sizeof(T) * cap_);
        else std::free(p_);
        p_ = nullptr;
        cap_ = 0;
    }
    T* p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
    bool mapped_ = false;
};

struct Sub {
    int32_t n = 0;
    BigVec<int64_t> xadj;
    BigVec<int32_t> adj, adjw, vw;
    BigVec<int32_t> gid;  // local -> global vertex id
    int64_t total_w = 0;
};

// Runs f(t) for t in [0, T): t = 0 on the calling thread, the others on
// their own threads.
template <class F>
static void run_threads(int T, F&& f) {
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back([&f, t]() { f(t); });
    f(0);
    for (auto& x : th) x.join();
}

// Below this many adjacency entries per thread a whole-graph pass runs on
// fewer threads.
constexpr int64_t PAR_EDGES_PER_THREAD = 1 << 20;

// Splits the vertices of s into T contiguous ranges of about equal
// vertices + adjacency entries; bounds[t]..bounds[t+1] is range t.
static std::vector<int32_t> vertex_chunks(const Sub& s, int T) {
    std::vector<int32_t> bounds(T + 1, s.n);
    bounds[0] = 0;
    const int64_t total = (int64_t)s.n + s.xadj[s.n];
    for (int t = 1; t < T; t++) {
        const int64_t want = total / T * t;
        int32_t lo = bounds[t - 1], hi = s.n;  // first v with v + xadj[v] >= want
        while (lo < hi) {
            const int32_t mid = lo + (hi - lo) / 2;
            if ((int64_t)mid + s.xadj[mid] < want) lo = mid + 1;
            else hi = mid;
        }
        bounds[t] = lo;
    }
    return bounds;
}

// Threads worth using for a whole-graph pass over s, at most par.
static int pass_threads(const Sub& s, int par) {
    const int64_t m = (int64_t)s.adj.size();
    return (int)std::max<int64_t>(1, std::min<int64_t>(par, m / PAR_EDGES_PER_THREAD));
}

static inline uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Max-priority queue of (gain, vertex) with lazy invalidation; ties broken
// toward the smaller vertex id so the result is deterministic.
struct GainQueue {
    struct E {
        int64_t g;
        int32_t v;
        bool operator<(const E& o) const { return g < o.g || (g == o.g && v > o.v); }
    };
    std::priority_queue<E> q;
    void push(int64_t g, int32_t v) { q.push({g, v}); }
    void clear() { q = std::priority_queue<E>(); }
};

struct Bisector {
    const Sub& s;
    std::vector<int8_t> side;
    std::vector<int64_t> ext, inn;  // edge weight to the other side / same side
    int64_t w[2] = {0, 0};
    int64_t cut = 0;

    explicit Bisector(const Sub& sub) : s(sub) {}

    // ext, inn, side weights and cut from `side`, on up to par threads (each
    // vertex's sums are its own; the totals are integer sums).
    void compute(int par = 1) {
        int32_t n = s.n;
        ext.resize(n);
        inn.resize(n);
        const int T = pass_threads(s, par);
        const std::vector<int32_t> bounds = vertex_chunks(s, T);
        std::vector<int64_t> part_w(2 * T, 0), part_cut(T, 0);
        run_threads(T, [&](int t) {
            int64_t w0 = 0, w1 = 0, c = 0;
            for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) {
                const int8_t sv = side[v];
                (sv ? w1 : w0) += s.vw[v];
                int64_t e = 0, i = 0;
                for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                    if (side[s.adj[j]] != sv) e += s.adjw[j];
                    else i += s.adjw[j];
                }
                ext[v] = e;
                inn[v] = i;
                c += e;
            }
            part_w[2 * t] = w0;
            part_w[2 * t + 1] = w1;
            part_cut[t] = c;
        });
        w[0] = w[1] = 0;
        cut = 0;
        for (int t = 0; t < T; t++) {
            w[0] += part_w[2 * t];
            w[1] += part_w[2 * t + 1];
            cut += part_cut[t];
        }
        cut /= 2;
    }

    void move(int32_t v) {
        int8_t from = side[v], to = (int8_t)(1 - from);
        side[v] = to;
        w[from] -= s.vw[v];
        w[to] += s.vw[v];
        cut -= ext[v] - inn[v];
        std::swap(ext[v], inn[v]);
        for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
            int32_t u = s.adj[j];
            int64_t ew = s.adjw[j];
            if (side[u] == to) {
                ext[u] -= ew;
                inn[u] += ew;
            } else {
                ext[u] += ew;
                inn[u] -= ew;
            }
        }
    }

    // Greedy graph growing of side 0 from `seed` until it reaches target0.
    void grow(int32_t seed, double target0, double maxw0) {
        int32_t n = s.n;
        side.assign(n, 1);
        compute();
        std::vector<uint8_t> tried(n, 0);
        GainQueue pq;
        auto push_nb = [&](int32_t v) {
            for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                int32_t u = s.adj[j];
                if (side[u] == 1 && !tried[u]) pq.push(ext[u] - inn[u], u);
            }
        };
        int32_t next_free = 0;
        int32_t cur = seed;
        while (w[0] < target0) {
            if (cur < 0) {
                while (!pq.q.empty()) {
                    auto e = pq.q.top();
                    pq.q.pop();
                    if (side[e.v] != 1 || tried[e.v] || e.g != ext[e.v] - inn[e.v]) continue;
                    cur = e.v;
                    break;
                }
            }
            if (cur < 0) {  // frontier empty: start a new region at the next untried vertex
                while (next_free < n && (side[next_free] != 1 || tried[next_free])) next_free++;
                if (next_free >= n) break;
                cur = next_free;
            }
            tried[cur] = 1;
            if ((double)(w[0] + s.vw[cur]) <= maxw0 || w[0] == 0) {
                move(cur);
                push_nb(cur);
            }
            cur = -1;
        }
    }

    // Balance-constrained 2-way FM. Keeps the best prefix of each pass; a pass
    // ends after `limit` moves without improvement (0: max(50, n/10)).
    void fm(const double maxw[2], int passes, int32_t limit_in = 0) {
        int32_t n = s.n;
        std::vector<uint8_t> locked(n);
        std::vector<int32_t> moved;
        for (int pass = 0; pass < passes; pass++) {
            std::fill(locked.begin(), locked.end(), 0);
            GainQueue pq[2];
            for (int32_t v = 0; v < n; v++)
                if (ext[v] > 0 || w[side[v]] > maxw[side[v]]) pq[side[v]].push(ext[v] - inn[v], v);
            auto over = [&]() {
                double o = 0;
                for (int t = 0; t < 2; t++)
                    if ((double)w[t] > maxw[t]) o += (double)w[t] - maxw[t];
                return o;
            };
            int64_t start_cut = cut;
            double best_over = over();
            int64_t best_cut = cut;
            size_t best_len = 0;
            moved.clear();
            int32_t limit = limit_in > 0 ? limit_in : std::max<int32_t>(50, n / 10);
            int32_t since_best = 0;
            while (since_best < limit) {
                int from = -1;
                int32_t pick[2] = {-1, -1};
                for (int t = 0; t < 2; t++) {
                    while (!pq[t].q.empty()) {
                        auto e = pq[t].q.top();
                        if (locked[e.v] || side[e.v] != t || e.g != ext[e.v] - inn[e.v]) {
                            pq[t].q.pop();
                            continue;
                        }
                        pick[t] = e.v;
                        break;
                    }
                }
                if ((double)w[0] > maxw[0]) from = 0;
                else if ((double)w[1] > maxw[1]) from = 1;
                if (from >= 0) {
                    if (pick[from] < 0) break;
                } else {
                    bool ok[2];
                    for (int t = 0; t < 2; t++)
                        ok[t] = pick[t] >= 0 && (double)(w[1 - t] + s.vw[pick[t]]) <= maxw[1 - t];
                    if (!ok[0] && !ok[1]) break;
                    if (ok[0] && ok[1]) {
                        int64_t g0 = ext[pick[0]] - inn[pick[0]], g1 = ext[pick[1]] - inn[pick[1]];
                        from = (g0 > g1 || (g0 == g1 && w[0] >= w[1])) ? 0 : 1;
                    } else {
                        from = ok[0] ? 0 : 1;
                    }
                }
                int32_t v = pick[from];
                pq[from].q.pop();
                locked[v] = 1;
                move(v);
                moved.push_back(v);
                for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                    int32_t u = s.adj[j];
                    if (!locked[u]) pq[side[u]].push(ext[u] - inn[u], u);
                }
                double o = over();
                if (o < best_over || (o == best_over && cut < best_cut)) {
                    best_over = o;
                    best_cut = cut;
                    best_len = moved.size();
                    since_best = 0;
                } else {
                    since_best++;
                }
            }
            while (moved.size() > best_len) {
                move(moved.back());
                moved.pop_back();
            }
            if (cut >= start_cut && best_len == 0) break;
        }
    }
};

constexpr int JET_HOST_THREADS_MAX = 16;

// Persistent host workers for the small, independent tasks of the recursion
// (trials of a bisection, the halves of a split): a task costs a queue push
// and a wake-up instead of a thread creation. Workers spin briefly before
// sleeping, since the tasks of one partition come in bursts. A thread that
// waits for its tasks runs queued tasks meanwhile (any group's), so nested
// waits cannot starve the pool. Created at first use, never destroyed.
class TaskPool {
  public:
    static TaskPool& get() {
        static TaskPool* p = new TaskPool(JET_HOST_THREADS_MAX - 1);
        return *p;
    }
    void push(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back(std::move(f));
            queued_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_one();
    }
    // Runs one queued task on the calling thread; false when none is queued.
    bool run_one() {
        std::function<void()> f;
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (q_.empty()) return false;
            f = std::move(q_.front());
            q_.pop_front();
            queued_.fetch_sub(1, std::memory_order_relaxed);
        }
        f();
        return true;
    }

  private:
    explicit TaskPool(int workers) {
        for (int i = 0; i < workers; i++) std::thread([this]() { loop(); }).detach();
    }
    void loop() {
        for (;;) {
            for (int spin = 0; spin < 20000; spin++) {
                if (queued_.load(std::memory_order_acquire) > 0 && run_one()) spin = 0;
                else std::this_thread::yield();
            }
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [this]() { return !q_.empty(); });
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    std::atomic<int> queued_{0};
};

// Tasks on the pool, joined by wait().
class TaskGroup {
  public:
    template <class F>
    void run(F&& f) {
        pending_.fetch_add(1, std::memory_order_relaxed);
        TaskPool::get().push([this, f]() {
            f();
            pending_.fetch_sub(1, std::memory_order_release);
        });
    }
    void wait() {
        while (pending_.load(std::memory_order_acquire) > 0)
            if (!TaskPool::get().run_one()) std::this_thread::yield();
    }
    ~TaskGroup() { wait(); }

  private:
    std::atomic<int> pending_{0};
};
// Below this many adjacency entries a subproblem is not worth a thread for
// its induced subgraphs or its coarsening; trials of one bisection and the
// two halves of a recursion get threads from much smaller sizes, since their
// work per thread (whole trials, whole subtrees) is larger, and they run as
// pool tasks. On a 557-vertex, 7430-entry graph split 64 ways, 8 threads take
// 1.9 ms against 9.0 ms on one (2.6 ms with a thread created per task).
constexpr int64_t PAR_MIN_EDGES = 200000;
constexpr int64_t PAR_MIN_TRIAL_EDGES = 500;
constexpr int64_t PAR_MIN_RECURSE_EDGES = 100;

struct Trial {
    std::vector<int8_t> side;
    double over = 0;
    int64_t cut = 0;
};

static int32_t pseudo_peripheral(const Sub& s) {
    int32_t n = s.n, seedv = 0;
    for (int rep = 0; rep < 2; rep++) {  // BFS twice from vertex 0
        std::vector<int32_t> dist(n, -1);
        std::vector<int32_t> qv;
        qv.push_back(seedv);
        dist[seedv] = 0;
        for (size_t h = 0; h < qv.size(); h++) {
            int32_t v = qv[h];
            for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++)
                if (dist[s.adj[j]] < 0) {
                    dist[s.adj[j]] = dist[v] + 1;
                    qv.push_back(s.adj[j]);
                }
        }
        seedv = qv.back();
    }
    return seedv;
}

static void run_trial(const Sub& s, int t, uint64_t seed, double target0, const double maxw[2], int passes,
                      Trial& out) {
    int32_t n = s.n;
    int32_t seedv = t == 0 ? pseudo_peripheral(s)
                           : (int32_t)(mix(seed + (uint64_t)t * 0x9E37ULL + (uint64_t)n) % (uint64_t)n);
    Bisector b(s);
    b.grow(seedv, target0, maxw[0]);
    b.fm(maxw, passes);
    out.over = 0;
    for (int side = 0; side < 2; side++)
        if ((double)b.w[side] > maxw[side]) out.over += (double)b.w[side] - maxw[side];
    out.cut = b.cut;
    out.side = std::move(b.side);
}

// Bisects `s` so side 0 carries about frac of its weight: several trials of
// greedy growing + FM on the whole graph, on up to `par` threads. Returns
// side per local vertex.
static std::vector<int8_t> bisect_flat(const Sub& s, double frac, double ub, uint64_t seed, int par) {
    double W = (double)s.total_w;
    double target0 = W * frac;
    double maxw[2] = {target0 * ub, (W - target0) * ub};
    int32_t n = s.n;
    int trials = n > 20000 ? 2 : (n > 2000 ? 4 : 8);
    int passes = n > 20000 ? 4 : 8;
    std::vector<Trial> res(trials);
    if (par > 1 && (int64_t)s.adj.size() >= PAR_MIN_TRIAL_EDGES) {
        // up to par runners (the caller and par - 1 pool tasks) take trials
        // in turn; each trial writes only its own result
        std::atomic<int> next{0};
        auto runner = [&]() {
            for (int t; (t = next.fetch_add(1)) < trials;) run_trial(s, t, seed, target0, maxw, passes, res[t]);
        };
        TaskGroup tg;
        for (int r = 1; r < std::min(par, trials); r++) tg.run(runner);
        runner();
        tg.wait();
    } else {
        for (int t = 0; t < trials; t++) run_trial(s, t, seed, target0, maxw, passes, res[t]);
    }
    int best = 0;
    for (int t = 1; t < trials; t++)
        if (res[t].over < res[best].over || (res[t].over == res[best].over && res[t].cut < res[best].cut)) best = t;
    return std::move(res[best].side);
}

// ---- multilevel bisection for large subgraphs (coarsen, bisect, refine up)

// Graphs above this size are coarsened before bisecting; coarsening stops at
// about ML_COARSEST vertices or when a level shrinks by less than 10%.
constexpr int32_t ML_MIN_N = 2000;
constexpr int32_t ML_COARSEST = 1000;
// graphs with at least this many adjacency entries are matched by
// handshakes on host threads instead of the serial greedy pass
constexpr int64_t PAR_MATCH_EDGES = int64_t(1) << 21;
constexpr int HANDSHAKE_ROUNDS = 6;

static void shuffle_ids(std::vector<int32_t>& p, uint64_t seed) {
    for (int32_t i = (int32_t)p.size() - 1; i > 0; i--) {
        int32_t j = (int32_t)(mix(seed + (uint64_t)i) % (uint64_t)(i + 1));
        std::swap(p[i], p[j]);
    }
}

// One coarsening level: heavy-edge matching under a vertex weight cap (in a
// seeded order, or by handshakes on large graphs), then pairs of
// still-unmatched vertices that share their first neighbour (leaves of one
// hub), then pairs of unmatched isolated vertices.
// cmap receives the coarse id of every vertex (ids in order of first member).
// Heavy-edge matching by handshakes, for large graphs, on up to par threads
// over contiguous vertex ranges: in each round every unmatched vertex
// proposes to its unmatched neighbour of heaviest edge (ties: larger seeded
// key) under the weight cap, and mutual proposals are matched. A vertex's
// proposal reads only the state at the start of the round, so the result does
// not depend on the thread count.
static void handshake_matching(const Sub& s, uint64_t seed, int64_t cap, int par, std::vector<int32_t>& mate) {
    const int32_t n = s.n;
    std::vector<int32_t> prop(n, -1);
    std::vector<uint32_t> key(n);
    const int T = pass_threads(s, par);
    const std::vector<int32_t> bounds = vertex_chunks(s, T);
    run_threads(T, [&](int t) {
        for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) key[v] = (uint32_t)mix(seed + (uint64_t)v);
    });
    for (int round = 0; round < HANDSHAKE_ROUNDS; round++) {
        std::vector<int64_t> proposed(T, 0);
        run_threads(T, [&](int t) {
            int64_t c = 0;
            for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) {
                prop[v] = -1;
                if (mate[v] >= 0) continue;
                int32_t best = -1;
                int64_t bw = -1;
                uint32_t bk = 0;
                for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                    const int32_t u = s.adj[j];
                    if (u == v || mate[u] >= 0 || (int64_t)s.vw[v] + s.vw[u] > cap) continue;
                    const int64_t w = s.adjw[j];
                    if (w > bw || (w == bw && key[u] > bk)) {
                        bw = w;
                        bk = key[u];
                        best = u;
                    }
                }
                prop[v] = best;
                c += best >= 0;
            }
            proposed[t] = c;
        });
        int64_t any = 0;
        for (int64_t c : proposed) any += c;
        if (any == 0) break;
        run_threads(T, [&](int t) {  // each vertex writes only its own mate
            for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) {
                const int32_t u = prop[v];
                if (u >= 0 && prop[u] == v) mate[v] = u;
            }
        });
    }
}

static Sub coarsen_once(const Sub& s, uint64_t seed, int par, std::vector<int32_t>& cmap) {
    const int32_t n = s.n;
    const int64_t cap = std::max<int64_t>(1, (int64_t)(1.5 * (double)s.total_w / ML_COARSEST));
    std::vector<int32_t> mate(n, -1);
    if ((int64_t)s.adj.size() >= PAR_MATCH_EDGES) {
        handshake_matching(s, seed, cap, par, mate);
    } else {
        std::vector<int32_t> order(n);
        for (int32_t i = 0; i < n; i++) order[i] = i;
        shuffle_ids(order, seed);
        for (int32_t v : order) {
            if (mate[v] >= 0) continue;
            int32_t best = -1;
            int64_t bw = -1;
            for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                int32_t u = s.adj[j];
                if (u == v || mate[u] >= 0 || (int64_t)s.vw[v] + s.vw[u] > cap) continue;
                if ((int64_t)s.adjw[j] > bw) {
                    bw = s.adjw[j];
                    best = u;
                }
            }
            if (best >= 0) {
                mate[v] = best;
                mate[best] = v;
            }
        }
    }
    // leaves sharing a neighbour, and isolated vertices, in increasing id
    std::vector<std::pair<int32_t, int32_t>> keyed;
    int32_t last_isolated = -1;
    for (int32_t v = 0; v < n; v++) {
        if (mate[v] >= 0) continue;
        if (s.xadj[v + 1] == s.xadj[v]) {
            if (last_isolated >= 0 && (int64_t)s.vw[v] + s.vw[last_isolated] <= cap) {
                mate[v] = last_isolated;
                mate[last_isolated] = v;
                last_isolated = -1;
            } else {
                last_isolated = v;
            }
            continue;
        }
        keyed.push_back({s.adj[s.xadj[v]], v});
    }
    std::sort(keyed.begin(), keyed.end());
    for (size_t i = 0; i + 1 < keyed.size();) {
        int32_t a = keyed[i].second, b = keyed[i + 1].second;
        if (keyed[i].first == keyed[i + 1].first && (int64_t)s.vw[a] + s.vw[b] <= cap) {
            mate[a] = b;
            mate[b] = a;
            i += 2;
        } else {
            i += 1;
        }
    }
    cmap.assign(n, -1);
    int32_t nc = 0;
    for (int32_t v = 0; v < n; v++) {
        if (cmap[v] >= 0) continue;
        cmap[v] = nc;
        if (mate[v] >= 0) cmap[mate[v]] = nc;
        nc++;
    }
    Sub c;
    c.n = nc;
    c.vw.assign(nc, 0);
    c.gid.resize(nc);
    std::vector<int32_t> first(nc, -1), second(nc, -1);
    for (int32_t v = 0; v < n; v++) {
        int32_t cv = cmap[v];
        c.vw[cv] += s.vw[v];
        if (first[cv] < 0) first[cv] = v;
        else second[cv] = v;
    }
    c.total_w = s.total_w;
    for (int32_t cv = 0; cv < nc; cv++) c.gid[cv] = cv;
    // Rows are built in contiguous chunks of coarse ids, one per thread, each
    // with its own position table; the chunks are then concatenated in order,
    // so the result is the same for any thread count.
    int T = std::max(1, std::min<int>(par, (int)(s.adj.size() / PAR_MIN_EDGES) + 1));
    std::vector<int32_t> bounds(T + 1, nc);
    {
        const int64_t per = ((int64_t)s.adj.size() + T - 1) / T;
        int64_t acc = 0;
        int t = 1;
        bounds[0] = 0;
        for (int32_t cv = 0; cv < nc && t < T; cv++) {
            for (int32_t v : {first[cv], second[cv]})
                if (v >= 0) acc += s.xadj[v + 1] - s.xadj[v];
            if (acc >= per * t) bounds[t++] = cv + 1;
        }
        for (; t < T; t++) bounds[t] = nc;
    }
    // pass 1 counts each coarse row's distinct neighbours, pass 2 writes the
    // rows at their final offsets; a row's entries keep first-seen order
    std::vector<int64_t> rowlen(nc, 0);
    auto pass = [&](int t, bool write) {
        std::vector<int64_t> slot(nc, -1);  // position of cu in the current row, if added there
        for (int32_t cv = bounds[t]; cv < bounds[t + 1]; cv++) {
            const int64_t row0 = write ? c.xadj[cv] : 0;
            int64_t len = 0;
            int64_t tag = write ? row0 : -2 - (int64_t)cv;  // count pass: slot holds -2 - row id
            for (int32_t v : {first[cv], second[cv]}) {
                if (v < 0) continue;
                for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                    int32_t cu = cmap[s.adj[j]];
                    if (cu == cv) continue;
                    if (write) {
                        if (slot[cu] < row0) {
                            slot[cu] = row0 + len++;
                            c.adj[slot[cu]] = cu;
                            c.adjw[slot[cu]] = s.adjw[j];
                        } else {
                            c.adjw[slot[cu]] += s.adjw[j];
                        }
                    } else if (slot[cu] != tag) {
                        slot[cu] = tag;
                        len++;
                    }
                }
            }
            if (!write) rowlen[cv] = len;
        }
    };
    auto run_all = [&](bool write) {
        std::vector<std::thread> th;
        for (int t = 1; t < T; t++) th.emplace_back(pass, t, write);
        pass(0, write);
        for (auto& x : th) x.join();
    };
    run_all(false);
    c.xadj.assign(nc + 1, 0);
    for (int32_t cv = 0; cv < nc; cv++) c.xadj[cv + 1] = c.xadj[cv] + rowlen[cv];
    c.adj.resize(c.xadj[nc]);
    c.adjw.resize(c.xadj[nc]);
    run_all(true);
    return c;
}

static std::vector<int8_t> bisect_flat(const Sub& s, double frac, double ub, uint64_t seed, int par);

// Coarsens s, bisects the coarsest graph with the flat method, then projects
// level by level with boundary FM passes.
static std::vector<int8_t> bisect_multilevel(const Sub& s, double frac, double ub, uint64_t seed, int par) {
    std::vector<Sub> levels;
    std::vector<std::vector<int32_t>> maps;
    const Sub* cur = &s;
    for (int lvl = 0; cur->n > ML_COARSEST && lvl < 64; lvl++) {
        std::vector<int32_t> cmap;
        Sub c = coarsen_once(*cur, seed + 0x51ED27ULL * (uint64_t)(lvl + 1), par, cmap);
        if ((double)c.n > 0.9 * (double)cur->n) break;
        maps.push_back(std::move(cmap));
        levels.push_back(std::move(c));
        cur = &levels.back();
    }
    std::vector<int8_t> side = bisect_flat(*cur, frac, ub, seed, par);
    const double W = (double)s.total_w;
    const double target0 = W * frac;
    const double maxw[2] = {target0 * ub, (W - target0) * ub};
    for (int lvl = (int)levels.size() - 1; lvl >= 0; lvl--) {
        const Sub& fine = lvl == 0 ? s : levels[lvl - 1];
        std::vector<int8_t> fs(fine.n);
        for (int32_t v = 0; v < fine.n; v++) fs[v] = side[maps[lvl][v]];
        Bisector b(fine);
        b.side = std::move(fs);
        b.compute(par);
        b.fm(maxw, 4, std::max<int32_t>(25, std::min<int32_t>(200, fine.n / 100)));
        side = std::move(b.side);
    }
    return side;
}

// The subgraph induced by the vertices with side == which, local ids in
// increasing order, on up to par threads over contiguous vertex ranges.
static Sub induced(const Sub& s, const std::vector<int8_t>& side, int8_t which, int par) {
    Sub r;
    const int T = pass_threads(s, par);
    const std::vector<int32_t> bounds = vertex_chunks(s, T);
    std::vector<int32_t> loc(s.n);
    std::vector<int32_t> first_loc(T + 1, 0);
    run_threads(T, [&](int t) {  // kept vertices per range
        int32_t c = 0;
        for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) c += side[v] == which;
        first_loc[t + 1] = c;
    });
    for (int t = 0; t < T; t++) first_loc[t + 1] += first_loc[t];
    const int32_t nr = first_loc[T];
    r.n = nr;
    r.gid.resize(nr);
    r.vw.resize(nr);
    r.xadj.resize((size_t)nr + 1);
    r.xadj[0] = 0;
    std::vector<int64_t> part_w(T, 0);
    run_threads(T, [&](int t) {  // local ids, then kept edges per kept vertex
        int32_t at = first_loc[t];
        for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) loc[v] = side[v] == which ? at++ : -1;
    });
    run_threads(T, [&](int t) {
        int64_t wsum = 0;
        for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) {
            if (loc[v] < 0) continue;
            int64_t c = 0;
            for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) c += side[s.adj[j]] == which;
            r.xadj[loc[v] + 1] = c;
            r.gid[loc[v]] = s.gid[v];
            r.vw[loc[v]] = s.vw[v];
            wsum += s.vw[v];
        }
        part_w[t] = wsum;
    });
    for (int t = 0; t < T; t++) r.total_w += part_w[t];
    for (int32_t i = 0; i < nr; i++) r.xadj[i + 1] += r.xadj[i];
    r.adj.resize(r.xadj[nr]);
    r.adjw.resize(r.xadj[nr]);
    run_threads(T, [&](int t) {  // fill
        for (int32_t v = bounds[t]; v < bounds[t + 1]; v++) {
            if (loc[v] < 0) continue;
            int64_t o = r.xadj[loc[v]];
            for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                const int32_t u = loc[s.adj[j]];
                if (u >= 0) {
                    r.adj[o] = u;
                    r.adjw[o] = s.adjw[j];
                    o++;
                }
            }
        }
    });
    return r;
}

static void recurse(const Sub& s, int32_t k, int32_t first, double ub, uint64_t seed, int32_t* part, int par) {
    if (k <= 1 || s.n == 0) {
        for (int32_t v = 0; v < s.n; v++) part[s.gid[v]] = first;
        return;
    }
    int32_t k0 = k / 2, k1 = k - k0;
    const double frac = (double)k0 / (double)k;
    const uint64_t bseed = seed + (uint64_t)first * 1315423911ULL;
    std::vector<int8_t> side = s.n > ML_MIN_N ? bisect_multilevel(s, frac, ub, bseed, par)
                                              : bisect_flat(s, frac, ub, bseed, par);
    Sub a = induced(s, side, 0, par);
    Sub b = induced(s, side, 1, par);
    side.clear();
    side.shrink_to_fit();
    if (par > 1 && (int64_t)(a.adj.size() + b.adj.size()) >= PAR_MIN_RECURSE_EDGES) {
        int pa = par / 2;
        TaskGroup tg;
        tg.run([&]() { recurse(a, k0, first, ub, seed, part, pa); });
        recurse(b, k1, first + k0, ub, seed, part, par - pa);
        tg.wait();
    } else {
        recurse(a, k0, first, ub, seed, part, par);
        recurse(b, k1, first + k0, ub, seed, part, par);
    }
}

// Threads of this host thread's partitions (0: no cap), set by the batched
// layer so that concurrent jobs share the host's CPUs.
thread_local int t_thread_budget = 0;

static int host_threads() {
    // the CPUs this thread may run on (a pinned process gets no more threads
    // than it has CPUs), else the machine's count
    int hw = 0;
    cpu_set_t cs;
    if (sched_getaffinity(0, sizeof(cs), &cs) == 0) hw = CPU_COUNT(&cs);
    if (hw < 1) hw = (int)std::thread::hardware_concurrency();
    if (hw < 1) hw = 1;
    if (const char* e = std::getenv("JET_HOST_THREADS")) {
        int v = std::atoi(e);
        if (v >= 1) hw = v;
    }
    if (t_thread_budget > 0) hw = std::min(hw, t_thread_budget);
    return std::min(hw, JET_HOST_THREADS_MAX);
}

}  // namespace

namespace jet_init {

void set_thread_budget(int threads) { t_thread_budget = threads > 0 ? threads : 0; }

std::vector<HostLevel> init_hierarchy(int32_t n, const int64_t* xadj, const int32_t* adj, const int32_t* adjw,
                                      const int32_t* vw, int32_t k, uint64_t seed) {
    std::vector<HostLevel> out;
    if (n <= 0) return out;
    Sub cur;
    cur.n = n;
    cur.xadj.assign(xadj, xadj + n + 1);
    const int64_t m = xadj[n];
    cur.adj.assign(adj, adj + m);
    if (adjw) cur.adjw.assign(adjw, adjw + m);
    else cur.adjw.assign(m, 1);
    if (vw) cur.vw.assign(vw, vw + n);
    else cur.vw.assign(n, 1);
    cur.gid.resize(n);
    for (int32_t v = 0; v < n; v++) {
        cur.gid[v] = v;
        cur.total_w += cur.vw[v];
    }
    const int32_t target = std::max<int32_t>(ML_COARSEST, 8 * k);
    const int par = host_threads();
    for (int lvl = 0; cur.n > target && lvl < 64; lvl++) {
        std::vector<int32_t> cmap;
        Sub c = coarsen_once(cur, seed + 0x1A17ULL * (uint64_t)(lvl + 1), par, cmap);
        if ((double)c.n > 0.9 * (double)cur.n) break;
        HostLevel L;
        L.n = c.n;
        L.xadj.assign(c.xadj.begin(), c.xadj.end());
        L.adj.assign(c.adj.begin(), c.adj.end());
        L.adjw.assign(c.adjw.begin(), c.adjw.end());
        L.vw.assign(c.vw.begin(), c.vw.end());
        L.map = std::move(cmap);
        out.push_back(std::move(L));
        cur = std::move(c);
    }
    cur = Sub();
    block_pool().trim();
    return out;
}

}  // namespace jet_init

// Returns 0 on success, -1 on invalid arguments.
extern "C" int jet_host_initpart(int32_t n, const int64_t* xadj, const int32_t* adj, const int32_t* adjw,
                                 const int32_t* vw, int32_t k, double imb, uint64_t seed, int32_t* part) {
    if (n < 0 || k < 1 || !(imb >= 1.0)) return -1;
    if (n == 0) return 0;
    Sub s;
    s.n = n;
    s.xadj.assign(xadj, xadj + n + 1);
    int64_t m = xadj[n];
    s.adj.assign(adj, adj + m);
    if (adjw) s.adjw.assign(adjw, adjw + m);
    else s.adjw.assign(m, 1);
    if (vw) s.vw.assign(vw, vw + n);
    else s.vw.assign(n, 1);
    s.gid.resize(n);
    for (int32_t v = 0; v < n; v++) {
        s.gid[v] = v;
        s.total_w += s.vw[v];
    }
    int depth = 0;
    while ((1 << depth) < k) depth++;
    double ub = depth > 0 ? std::pow(imb, 1.0 / depth) : imb;
    recurse(s, k, 0, ub, seed, part, host_threads());
    s = Sub();
    block_pool().trim();
    return 0;
}
