// Calls all nine jet_partitioner entry points of include/jet.h on one 2-D
// grid with the same config and seed, and prints one line per call:
// "<name> <cut> <checksum>". Exit status: 0 when every call returned the
// same partition, 1 when they differ, 3 when the library raised an error
// (the line "error <code> <message>" names it).
#include <algorithm>
#include <cstdio>
#include <vector>

#include "jet.h"

namespace jp = jet_partitioner;

template <class T>
static T* dev_copy(const std::vector<T>& h) {
    T* d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(T) * h.size()) != hipSuccess) throw jp::error(JET_ENOMEM, "hipMalloc");
    if (hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess) throw jp::error(JET_EDEVICE, "hipMemcpy");
    return d;
}

static unsigned long long checksum(const std::vector<int>& p) {
    unsigned long long h = 1469598103934665603ULL;
    for (int v : p) h = (h ^ (unsigned)v) * 1099511628211ULL;
    return h;
}

int main() {
    const int nx = 96, ny = 64, n = nx * ny;
    std::vector<int32_t> rp32(1, 0), col32, ew32;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const int nb[4][2] = {{x, y - 1}, {x - 1, y}, {x + 1, y}, {x, y + 1}};
            for (auto& q : nb)
                if (q[0] >= 0 && q[0] < nx && q[1] >= 0 && q[1] < ny) {
                    const int u = y * nx + x, v = q[1] * nx + q[0];
                    col32.push_back(v);
                    ew32.push_back(1 + (7 * std::min(u, v) + std::max(u, v)) % 4);  // symmetric
                }
            rp32.push_back((int32_t)col32.size());
        }
    std::vector<int64_t> rp64(rp32.begin(), rp32.end()), col64(col32.begin(), col32.end()),
        ew64(ew32.begin(), ew32.end());
    std::vector<int32_t> vw32(n, 1);
    std::vector<int64_t> vw64(n, 1);

    jp::config_t cfg;
    cfg.num_parts = 8;
    cfg.coarsening_alg = 1;
    const uint64_t seed = 5;
    std::vector<std::vector<int>> parts;
    try {
        {
            jp::host_matrix_t g{n, rp32.data(), col32.data(), ew32.data()};
            jp::wgt_host_vt w{vw32.data(), n};
            jp::experiment_data<jp::value_t> ex;
            jp::value_t cut = 0;
            parts.push_back(jp::partition_host(cut, cfg, g, w, false, ex, seed));
            std::printf("partition_host %d %llu\n", cut, checksum(parts.back()));
            parts.push_back(jp::partition_serial(cut, cfg, g, w, false, ex, seed));
            std::printf("partition_serial %d %llu\n", cut, checksum(parts.back()));
        }
        {
            jp::big_host_matrix_t g{n, rp64.data(), col32.data(), ew64.data()};
            jp::big_wgt_host_vt w{vw64.data(), n};
            jp::experiment_data<jp::big_val_t> ex;
            jp::big_val_t cut = 0;
            parts.push_back(jp::partition_host_big(cut, cfg, g, w, false, ex, seed));
            std::printf("partition_host_big %lld %llu\n", (long long)cut, checksum(parts.back()));
            parts.push_back(jp::partition_serial_big(cut, cfg, g, w, false, ex, seed));
            std::printf("partition_serial_big %lld %llu\n", (long long)cut, checksum(parts.back()));
        }
        {
            jp::biggest_host_matrix_t g{n, rp64.data(), col64.data(), ew64.data()};
            jp::big_wgt_host_vt w{vw64.data(), n};
            jp::experiment_data<jp::big_val_t> ex;
            jp::big_val_t cut = 0;
            parts.push_back(jp::partition_host_biggest(cut, cfg, g, w, false, ex, seed));
            std::printf("partition_host_biggest %lld %llu\n", (long long)cut, checksum(parts.back()));
            parts.push_back(jp::partition_serial_biggest(cut, cfg, g, w, false, ex, seed));
            std::printf("partition_serial_biggest %lld %llu\n", (long long)cut, checksum(parts.back()));
        }
        {
            jp::matrix_t g{n, dev_copy(rp32), dev_copy(col32), dev_copy(ew32)};
            jp::wgt_vt w{dev_copy(vw32), n};
            jp::experiment_data<jp::value_t> ex;
            jp::value_t cut = 0;
            parts.push_back(jp::partition(cut, cfg, g, w, false, ex, seed).to_host());
            std::printf("partition %d %llu levels %d\n", cut, checksum(parts.back()), ex.num_levels());
        }
        {
            jp::big_matrix_t g{n, dev_copy(rp64), dev_copy(col32), dev_copy(ew64)};
            jp::big_wgt_vt w{dev_copy(vw64), n};
            jp::experiment_data<jp::big_val_t> ex;
            jp::big_val_t cut = 0;
            parts.push_back(jp::partition_big(cut, cfg, g, w, false, ex, seed).to_host());
            std::printf("partition_big %lld %llu\n", (long long)cut, checksum(parts.back()));
        }
        {
            jp::biggest_matrix_t g{n, dev_copy(rp64), dev_copy(col64), dev_copy(ew64)};
            jp::big_wgt_vt w{dev_copy(vw64), n};
            jp::experiment_data<jp::big_val_t> ex;
            jp::big_val_t cut = 0;
            parts.push_back(jp::partition_biggest(cut, cfg, g, w, false, ex, seed).to_host());
            std::printf("partition_biggest %lld %llu\n", (long long)cut, checksum(parts.back()));
        }
    } catch (const jp::error& e) {
        std::printf("error %d %s\n", e.code, e.what());
        return 3;
    }
    for (auto& p : parts)
        if (p != parts[0]) return 1;
    return 0;
}
