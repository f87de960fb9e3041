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
// start refinement from the same coarsest partition.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <queue>
#include <vector>

namespace {

struct Sub {
    int32_t n = 0;
    std::vector<int64_t> xadj;
    std::vector<int32_t> adj, adjw, vw;
    std::vector<int32_t> gid;  // local -> global vertex id
    int64_t total_w = 0;
};

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

    void compute() {
        int32_t n = s.n;
        ext.assign(n, 0);
        inn.assign(n, 0);
        w[0] = w[1] = 0;
        cut = 0;
        for (int32_t v = 0; v < n; v++) {
            w[side[v]] += s.vw[v];
            for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
                if (side[s.adj[j]] != side[v]) ext[v] += s.adjw[j];
                else inn[v] += s.adjw[j];
            }
            cut += ext[v];
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

    // Balance-constrained 2-way FM. Keeps the best prefix of each pass.
    void fm(const double maxw[2], int passes) {
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
            int32_t limit = std::max<int32_t>(50, n / 10);
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

// Bisects `s` so side 0 carries about frac of its weight. Returns side per
// local vertex.
static std::vector<int8_t> bisect(const Sub& s, double frac, double ub, uint64_t seed) {
    double W = (double)s.total_w;
    double target0 = W * frac;
    double maxw[2] = {target0 * ub, (W - target0) * ub};
    int32_t n = s.n;
    int trials = n > 20000 ? 2 : (n > 2000 ? 4 : 8);
    int passes = n > 20000 ? 4 : 8;
    std::vector<int8_t> best;
    double best_over = 0;
    int64_t best_cut = 0;
    Bisector b(s);
    for (int t = 0; t < trials; t++) {
        int32_t seedv;
        if (t == 0) {  // pseudo-peripheral vertex: BFS twice from vertex 0
            seedv = 0;
            for (int rep = 0; rep < 2; rep++) {
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
        } else {
            seedv = (int32_t)(mix(seed + (uint64_t)t * 0x9E37ULL + (uint64_t)n) % (uint64_t)n);
        }
        b.grow(seedv, target0, maxw[0]);
        b.fm(maxw, passes);
        double over = 0;
        for (int side = 0; side < 2; side++)
            if ((double)b.w[side] > maxw[side]) over += (double)b.w[side] - maxw[side];
        if (best.empty() || over < best_over || (over == best_over && b.cut < best_cut)) {
            best = b.side;
            best_over = over;
            best_cut = b.cut;
        }
    }
    return best;
}

static Sub induced(const Sub& s, const std::vector<int8_t>& side, int8_t which) {
    Sub r;
    std::vector<int32_t> loc(s.n, -1);
    for (int32_t v = 0; v < s.n; v++)
        if (side[v] == which) {
            loc[v] = r.n++;
            r.gid.push_back(s.gid[v]);
            r.vw.push_back(s.vw[v]);
            r.total_w += s.vw[v];
        }
    r.xadj.assign(1, 0);
    for (int32_t v = 0; v < s.n; v++) {
        if (side[v] != which) continue;
        for (int64_t j = s.xadj[v]; j < s.xadj[v + 1]; j++) {
            int32_t u = s.adj[j];
            if (loc[u] >= 0) {
                r.adj.push_back(loc[u]);
                r.adjw.push_back(s.adjw[j]);
            }
        }
        r.xadj.push_back((int64_t)r.adj.size());
    }
    return r;
}

static void recurse(const Sub& s, int32_t k, int32_t first, double ub, uint64_t seed, int32_t* part) {
    if (k <= 1 || s.n == 0) {
        for (int32_t v = 0; v < s.n; v++) part[s.gid[v]] = first;
        return;
    }
    int32_t k0 = k / 2, k1 = k - k0;
    std::vector<int8_t> side = bisect(s, (double)k0 / (double)k, ub, seed + (uint64_t)first * 1315423911ULL);
    Sub a = induced(s, side, 0);
    Sub b = induced(s, side, 1);
    recurse(a, k0, first, ub, seed, part);
    recurse(b, k1, first + k0, ub, seed, part);
}

}  // namespace

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
    recurse(s, k, 0, ub, seed, part);
    return 0;
}
