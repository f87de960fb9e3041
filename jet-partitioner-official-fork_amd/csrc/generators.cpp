// Synthetic graph generators for the measured workloads (host).
// Output: symmetric CSR, rows sorted, no self-loops, no duplicates.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "jet_capi.h"

namespace {

inline uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" int jet_generate_grid2d(int64_t nx, int64_t ny, int64_t* rowptr, int32_t* col) {
    if (nx < 1 || ny < 1 || nx * ny >= (int64_t(1) << 31)) return JET_EINVAL;
    int64_t e = 0;
    for (int64_t y = 0; y < ny; y++) {
        for (int64_t x = 0; x < nx; x++) {
            int64_t v = y * nx + x;
            rowptr[v] = e;
            if (y > 0) col[e++] = (int32_t)(v - nx);
            if (x > 0) col[e++] = (int32_t)(v - 1);
            if (x + 1 < nx) col[e++] = (int32_t)(v + 1);
            if (y + 1 < ny) col[e++] = (int32_t)(v + nx);
        }
    }
    rowptr[nx * ny] = e;
    return JET_OK;
}

extern "C" int jet_generate_stencil3d(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, int32_t wmax, int64_t* rowptr,
                                      int32_t* col, int32_t* ew) {
    int64_t N = nx * ny * nz;
    if (nx < 1 || ny < 1 || nz < 1 || N >= (int64_t(1) << 31)) return JET_EINVAL;
    int64_t e = 0;
    const int64_t sxy = nx * ny;
    for (int64_t z = 0; z < nz; z++)
        for (int64_t y = 0; y < ny; y++)
            for (int64_t x = 0; x < nx; x++) {
                int64_t v = (z * ny + y) * nx + x;
                rowptr[v] = e;
                int64_t nb[6];
                int c = 0;
                if (z > 0) nb[c++] = v - sxy;
                if (y > 0) nb[c++] = v - nx;
                if (x > 0) nb[c++] = v - 1;
                if (x + 1 < nx) nb[c++] = v + 1;
                if (y + 1 < ny) nb[c++] = v + nx;
                if (z + 1 < nz) nb[c++] = v + sxy;
                for (int t = 0; t < c; t++) {
                    int64_t u = nb[t];
                    col[e] = (int32_t)u;
                    if (ew) {
                        uint64_t a = (uint64_t)std::min(u, v), b = (uint64_t)std::max(u, v);
                        ew[e] = wmax > 0 ? (int32_t)(1 + splitmix(seed ^ (a * (uint64_t)N + b)) % (uint64_t)wmax) : 1;
                    }
                    e++;
                }
            }
    rowptr[N] = e;
    return JET_OK;
}

extern "C" int jet_generate_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed,
                                 int64_t* rowptr, int32_t* col, int64_t* m_out) {
    if (scale < 1 || scale > 30 || ef < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return JET_EINVAL;
    const int64_t n = int64_t(1) << scale;
    const int64_t draws = (int64_t)ef * n;
    // one 64-bit key (min << 32 | max) per drawn undirected edge
    std::vector<uint64_t> keys;
    keys.reserve(draws);
    const double ab = a + b, abc = a + b + c;
    uint64_t state = splitmix(seed ^ 0x5DEECE66DULL);
    for (int64_t i = 0; i < draws; i++) {
        uint64_t u = 0, v = 0;
        for (int l = 0; l < scale; l++) {
            state = splitmix(state);
            double r = (double)(state >> 11) * (1.0 / 9007199254740992.0);
            uint64_t bu = r >= ab ? 1 : 0;
            uint64_t bv = ((r >= a && r < ab) || r >= abc) ? 1 : 0;
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        if (u == v) continue;
        uint64_t lo = std::min(u, v), hi = std::max(u, v);
        keys.push_back((lo << 32) | hi);
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<int64_t> deg(n + 1, 0);
    for (uint64_t k : keys) {
        deg[(k >> 32) + 1]++;
        deg[(k & 0xffffffffULL) + 1]++;
    }
    for (int64_t i = 0; i < n; i++) deg[i + 1] += deg[i];
    for (int64_t i = 0; i <= n; i++) rowptr[i] = deg[i];
    std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
    for (uint64_t k : keys) {
        int64_t lo = (int64_t)(k >> 32), hi = (int64_t)(k & 0xffffffffULL);
        col[fill[lo]++] = (int32_t)hi;
        col[fill[hi]++] = (int32_t)lo;
    }
    for (int64_t i = 0; i < n; i++) std::sort(col + rowptr[i], col + rowptr[i + 1]);
    *m_out = rowptr[n];
    return JET_OK;
}
