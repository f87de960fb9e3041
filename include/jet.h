/*
 * C++ mirror of the reference's partition entry points (header/jet.h:46-107,
 * namespace jet_partitioner) over the C-ABI in jet_capi.h.
 *
 * The nine names, their parameter order and config_t's fields and defaults
 * (header/jet_config.h:43-52) are kept. Kokkos views are replaced by plain
 * CSR handles: crs_matrix<...> holds row_map / entries / values pointers
 * plus the row count, tagged with the memory space the pointers live in.
 * Every variant runs on the GPU. The _host and _serial names take host arrays
 * and return the partition as a host vector (part_mt), as the reference's host
 * variants do; the plain names take device arrays and return a device array
 * (part_vt). Errors throw jet_partitioner::error carrying the C-ABI code; the
 * reference has no error path (METIS failure exits, initial_partition.hpp:106-109).
 * An optional trailing seed replaces the reference's std::time seed
 * (contract.hpp:503).
 *
 * Header-only; link against libjet_amd.so and the HIP runtime (amdhip64).
 */
#ifndef JET_PARTITIONER_JET_H
#define JET_PARTITIONER_JET_H

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "jet_capi.h"

namespace jet_partitioner {

/* header/jet_config.h:43-52 */
struct config_t {
    int coarsening_alg = 0;
    int num_iter = 1;
    double max_imb_ratio = 1.03;
    int num_parts = 2;
    double refine_tolerance = 0.999;
    bool dump_coarse = false;
    bool verbose = false;
    bool ultra_settings = false;
};

/* header/jet_defs.h:47-53,64 */
using big_ordinal_t = int64_t;
using big_offset_t = int64_t;
using big_val_t = int64_t;
using ordinal_t = int32_t;
using edge_offset_t = int32_t;
using value_t = int32_t;
using part_t = int;

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

struct device_space {};
struct host_space {};

/* Stand-in for KokkosSparse::CrsMatrix: symmetric CSR, both directions stored.
 * values may be null (unit edge weights). */
template <class Value, class Ordinal, class Space, class Offset>
struct crs_matrix {
    Ordinal num_rows = 0;
    const Offset* row_map = nullptr;  // num_rows + 1
    const Ordinal* entries = nullptr;  // row_map[num_rows]
    const Value* values = nullptr;     // row_map[num_rows], or null
    Ordinal numRows() const { return num_rows; }
};

/* Stand-in for a Kokkos::View<T*> of vertex weights (may be null: unit). */
template <class T, class Space>
struct weight_view {
    const T* ptr = nullptr;
    int64_t n = 0;
    const T* data() const { return ptr; }
    int64_t extent(int) const { return n; }
};

/* Owned device array of parts (part_vt, header/jet_defs.h:69). */
class device_parts {
  public:
    device_parts() = default;
    explicit device_parts(int64_t n) : n_(n) {
        part_t* p = nullptr;
        if (n > 0 && hipMalloc(reinterpret_cast<void**>(&p), sizeof(part_t) * (size_t)n) != hipSuccess)
            throw error(JET_ENOMEM, "hipMalloc failed for the partition");
        p_.reset(p, [](part_t* q) {
            if (q) (void)hipFree(q);
        });
    }
    part_t* data() const { return p_.get(); }
    int64_t extent(int) const { return n_; }
    int64_t size() const { return n_; }
    /* Copies the partition to the host. */
    std::vector<part_t> to_host() const {
        std::vector<part_t> h((size_t)n_);
        if (n_ > 0 && hipMemcpy(h.data(), p_.get(), sizeof(part_t) * (size_t)n_, hipMemcpyDeviceToHost) != hipSuccess)
            throw error(JET_EDEVICE, "hipMemcpy failed");
        return h;
    }

  private:
    std::shared_ptr<part_t> p_;
    int64_t n_ = 0;
};

using part_vt = device_parts;
using part_mt = std::vector<part_t>;

using matrix_t = crs_matrix<value_t, ordinal_t, device_space, edge_offset_t>;
using host_matrix_t = crs_matrix<value_t, ordinal_t, host_space, edge_offset_t>;
using serial_matrix_t = host_matrix_t;
using wgt_vt = weight_view<value_t, device_space>;
using wgt_host_vt = weight_view<value_t, host_space>;
using wgt_serial_vt = wgt_host_vt;

using big_matrix_t = crs_matrix<big_val_t, ordinal_t, device_space, big_offset_t>;
using big_host_matrix_t = crs_matrix<big_val_t, ordinal_t, host_space, big_offset_t>;
using big_serial_matrix_t = big_host_matrix_t;
using big_wgt_vt = weight_view<big_val_t, device_space>;
using big_wgt_host_vt = weight_view<big_val_t, host_space>;
using big_wgt_serial_vt = big_wgt_host_vt;

using biggest_matrix_t = crs_matrix<big_val_t, big_ordinal_t, device_space, big_offset_t>;
using biggest_host_matrix_t = crs_matrix<big_val_t, big_ordinal_t, host_space, big_offset_t>;
using biggest_serial_matrix_t = biggest_host_matrix_t;

/* experiment_data<T> (header/experiment_data.hpp:50-243): phase timers and
 * per-level refinement records, coarsest level first. */
template <class T>
struct experiment_data {
    jet_stats_t stats{};
    int32_t num_levels() const { return stats.num_levels; }
    const jet_level_stats_t& level(int32_t i) const { return stats.levels[i]; }
    double total_seconds() const { return stats.total_s; }
    T finest_edge_cut() const { return (T)stats.finest_edge_cut; }
    double finest_imb_ratio() const { return stats.finest_imb_ratio; }
};

namespace detail {

inline jet_config_t to_c(const config_t& c) {
    jet_config_t o;
    jet_config_default(&o);
    o.coarsening_alg = c.coarsening_alg;
    o.num_iter = c.num_iter;
    o.max_imb_ratio = c.max_imb_ratio;
    o.num_parts = c.num_parts;
    o.refine_tolerance = c.refine_tolerance;
    o.dump_coarse = c.dump_coarse ? 1 : 0;
    o.verbose = c.verbose ? 1 : 0;
    o.ultra_settings = c.ultra_settings ? 1 : 0;
    return o;
}

inline void check(int rc) {
    if (rc != JET_OK) throw error(rc, jet_last_error());
}

/* One call through the C-ABI for any width combination. */
template <class Value, class Ordinal, class Offset, class Space, class W, class Cut>
inline void run(Cut& edge_cut, const config_t& config, const crs_matrix<Value, Ordinal, Space, Offset>& g,
                const weight_view<W, Space>& vw, bool uniform_ew, experiment_data<Cut>& experiment, uint64_t seed,
                part_t* out) {
    const jet_config_t c = to_c(config);
    const int space = std::is_same<Space, device_space>::value ? JET_PTR_DEVICE : JET_PTR_HOST;
    int64_t cut = 0;
    int rc;
    if constexpr (sizeof(Ordinal) == 8) {
        rc = jet_partition_csr_biggest(&c, g.num_rows, g.row_map, g.entries, g.values, vw.ptr, uniform_ew, space, seed,
                                       out, &cut, &experiment.stats);
    } else if constexpr (sizeof(Value) == 8) {
        rc = jet_partition_csr_big(&c, g.num_rows, g.row_map, g.entries, g.values, vw.ptr, uniform_ew, space, seed, out,
                                   &cut, &experiment.stats);
    } else if constexpr (sizeof(Offset) == 8) {
        rc = jet_partition_csr(&c, g.num_rows, g.row_map, g.entries, g.values, vw.ptr, uniform_ew, space, seed, out,
                               &cut, &experiment.stats);
    } else {
        rc = jet_partition_csr32(&c, g.num_rows, g.row_map, g.entries, g.values, vw.ptr, uniform_ew, space, seed, out,
                                 &cut, &experiment.stats);
    }
    check(rc);
    edge_cut = (Cut)cut;
}

template <class M, class W, class Cut>
inline part_vt on_device(Cut& edge_cut, const config_t& config, const M& g, const W& vw, bool uniform_ew,
                         experiment_data<Cut>& experiment, uint64_t seed) {
    part_vt p(g.num_rows);
    run(edge_cut, config, g, vw, uniform_ew, experiment, seed, p.data());
    return p;
}

template <class M, class W, class Cut>
inline part_mt on_host(Cut& edge_cut, const config_t& config, const M& g, const W& vw, bool uniform_ew,
                       experiment_data<Cut>& experiment, uint64_t seed) {
    part_mt p((size_t)g.num_rows);
    run(edge_cut, config, g, vw, uniform_ew, experiment, seed, p.data());
    return p;
}

}  // namespace detail

/* header/jet.h:46-51; src/jet.cpp:47-55 */
inline part_vt partition(value_t& edge_cut, const config_t& config, const matrix_t g, const wgt_vt vweights,
                         bool uniform_ew, experiment_data<value_t>& experiment, uint64_t seed = 0) {
    return detail::on_device(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:53-58 */
inline part_mt partition_host(value_t& edge_cut, const config_t& config, const host_matrix_t g,
                              const wgt_host_vt vweights, bool uniform_ew, experiment_data<value_t>& experiment,
                              uint64_t seed = 0) {
    return detail::on_host(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:60-65 */
inline part_mt partition_serial(value_t& edge_cut, const config_t& config, const serial_matrix_t g,
                                const wgt_serial_vt vweights, bool uniform_ew, experiment_data<value_t>& experiment,
                                uint64_t seed = 0) {
    return detail::on_host(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:67-72; src/jet_big.cpp:47-55 */
inline part_vt partition_big(big_val_t& edge_cut, const config_t& config, const big_matrix_t g,
                             const big_wgt_vt vweights, bool uniform_ew, experiment_data<big_val_t>& experiment,
                             uint64_t seed = 0) {
    return detail::on_device(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:74-79 */
inline part_mt partition_host_big(big_val_t& edge_cut, const config_t& config, const big_host_matrix_t g,
                                  const big_wgt_host_vt vweights, bool uniform_ew,
                                  experiment_data<big_val_t>& experiment, uint64_t seed = 0) {
    return detail::on_host(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:81-86 */
inline part_mt partition_serial_big(big_val_t& edge_cut, const config_t& config, const big_serial_matrix_t g,
                                    const big_wgt_serial_vt vweights, bool uniform_ew,
                                    experiment_data<big_val_t>& experiment, uint64_t seed = 0) {
    return detail::on_host(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:88-93; src/jet_biggest.cpp:47-55 */
inline part_vt partition_biggest(big_val_t& edge_cut, const config_t& config, const biggest_matrix_t g,
                                 const big_wgt_vt vweights, bool uniform_ew, experiment_data<big_val_t>& experiment,
                                 uint64_t seed = 0) {
    return detail::on_device(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:95-100 */
inline part_mt partition_host_biggest(big_val_t& edge_cut, const config_t& config, const biggest_host_matrix_t g,
                                      const big_wgt_host_vt vweights, bool uniform_ew,
                                      experiment_data<big_val_t>& experiment, uint64_t seed = 0) {
    return detail::on_host(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

/* header/jet.h:102-107 */
inline part_mt partition_serial_biggest(big_val_t& edge_cut, const config_t& config, const biggest_serial_matrix_t g,
                                        const big_wgt_serial_vt vweights, bool uniform_ew,
                                        experiment_data<big_val_t>& experiment, uint64_t seed = 0) {
    return detail::on_host(edge_cut, config, g, vweights, uniform_ew, experiment, seed);
}

}  // namespace jet_partitioner

#endif /* JET_PARTITIONER_JET_H */
