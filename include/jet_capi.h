/*
 * C-ABI of the MI355X Jet partitioner (libjet_amd.so).
 *
 * The reference exposes C++ overloads over Kokkos views
 * (header/jet.h:46-107, namespace jet_partitioner); an FFI cannot bind those.
 * This header is the raw-CSR boundary that replaces them: plain pointers and
 * sizes, caller-owned output, integer error codes, no exceptions.
 *
 * Graph input: symmetric CSR with both directions stored (the reference's
 * unchecked precondition), row offsets monotone from 0, columns in [0, n),
 * positive weights. Edge/vertex weight arrays may be NULL (unit weights).
 * `ptr_space` says whether every array argument of the call lives in host
 * memory (JET_PTR_HOST) or HIP device memory (JET_PTR_DEVICE).
 */
#ifndef JET_CAPI_H
#define JET_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define JET_OK 0
#define JET_EAGAIN 1         /* jet_poll: the job has not finished yet */
#define JET_EINVAL (-1)      /* malformed CSR, bad config, k out of range */
#define JET_ENOMEM (-2)      /* device allocation failed */
#define JET_EINIT (-3)       /* initial partitioner failed (the reference exits, initial_partition.hpp:106-109) */
#define JET_EDEVICE (-4)     /* HIP error or no usable GPU */
#define JET_EUNSUPPORTED (-5)

#define JET_PTR_HOST 0
#define JET_PTR_DEVICE 1

#define JET_MAX_PARTS 8192
#define JET_MAX_LEVELS 256

/* config_t (header/jet_config.h:43-52): same fields, same meaning.
 * coarsening_alg: 0 = 2-hop matching (MtMetis), 1 = HEC, 2 = pure matching,
 * other = 2-hop matching (partitioner.hpp:75-87). num_iter: attempts, best
 * edge cut kept by strict < (app/driver.cpp:110-130). dump_coarse: write the
 * coarse hierarchy during the call (see jet_load_hierarchy below). */
typedef struct jet_config_t {
    int32_t coarsening_alg;
    int32_t num_iter;
    double max_imb_ratio;
    int32_t num_parts;
    double refine_tolerance;
    int32_t dump_coarse;
    int32_t verbose;
    int32_t ultra_settings;
} jet_config_t;

/* experiment_data::CoarseLevel (header/experiment_data.hpp:102-123). */
typedef struct jet_level_stats_t {
    int64_t edge_cut;
    double imb;
    int64_t num_edges;
    int64_t num_vertices;
    double total_ref_time;
    int32_t total_iterations;
    int32_t lp_iterations;
} jet_level_stats_t;

/* experiment_data<T> (header/experiment_data.hpp:50-243): phase timers in
 * seconds and per-level refinement records, coarsest level first. */
typedef struct jet_stats_t {
    double total_s;
    double coarsen_s;
    double map_s;
    double build_s;
    double init_partition_s;
    double refine_s;
    int32_t num_levels;
    int32_t attempts;
    double finest_imb_ratio;
    int64_t finest_edge_cut;
    int64_t largest_part;
    int64_t smallest_part;
    jet_level_stats_t levels[JET_MAX_LEVELS];
} jet_stats_t;

/* config_t defaults (header/jet_config.h:43-52). */
void jet_config_default(jet_config_t* cfg);

/* Message for the last non-zero return on this thread. */
const char* jet_last_error(void);

/* Number of visible HIP devices (0 when none). */
int jet_device_count(void);

/* Partition a graph into cfg->num_parts parts.
 * Replaces partition_big / partition_host_big / partition_serial_big
 * (header/jet.h:74-93; src/jet_big.cpp:47-75): int64 offsets, int32 columns,
 * here int32 weights. part_out (n entries, caller-owned, in ptr_space)
 * receives the best partition; *cut_out the undirected edge cut
 * (uncoarsen.hpp:136). seed replaces the reference's std::time seed
 * (contract.hpp:503); attempt i uses seed + i. stats may be NULL. */
int jet_partition_csr(const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int32_t* col,
                      const int32_t* ew, const int32_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
                      int32_t* part_out, int64_t* cut_out, jet_stats_t* stats);

/* Same with int32 row offsets: replaces partition / partition_host /
 * partition_serial (header/jet.h:46-72; src/jet.cpp:47-75). */
int jet_partition_csr32(const jet_config_t* cfg, int64_t n, const int32_t* row_offsets, const int32_t* col,
                        const int32_t* ew, const int32_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
                        int32_t* part_out, int64_t* cut_out, jet_stats_t* stats);

/* Same with int64 weights (big_val_t, header/jet_defs.h:48): values are
 * narrowed to int32 after a range check (total weight < 2^31). */
int jet_partition_csr_big(const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int32_t* col,
                          const int64_t* ew, const int64_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
                          int32_t* part_out, int64_t* cut_out, jet_stats_t* stats);

/* Same with int64 columns: replaces partition_biggest / partition_host_biggest
 * / partition_serial_biggest (header/jet.h:88-107; src/jet_biggest.cpp:47-75,
 * big_ordinal_t = int64, header/jet_defs.h:89-92). The device path keeps
 * 32-bit vertex ids, so n must be below 2^31 - 2 (else JET_EINVAL); columns
 * are narrowed on the device after a range check. */
int jet_partition_csr_biggest(const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int64_t* col,
                              const int64_t* ew, const int64_t* vw, int32_t uniform_ew, int32_t ptr_space,
                              uint64_t seed, int32_t* part_out, int64_t* cut_out, jet_stats_t* stats);

/* ---- Sub-function entry points (one hot-path stage each; host or device
 * pointers per ptr_space). They exist so integrators and parity tests can
 * drive one stage of the path on the GPU. */

/* contract.hpp:448-491 (generate_coarse_mapping) for one level: alg as
 * config_t::coarsening_alg, level = index of the level being coarsened
 * (0 = input graph), uniform = the level's uniform-edge-weight flag.
 * Writes map_out[n] and *nc_out. */
int jet_coarsen_map(int32_t alg, int32_t level, int32_t uniform, uint64_t seed, int64_t n, const int64_t* row_offsets,
                    const int32_t* col, const int32_t* ew, const int32_t* vw, int32_t ptr_space, int32_t* map_out,
                    int32_t* nc_out);

/* heuristics.hpp:195-248: HEC heavy-neighbour choice (hn_out[n]). */
int jet_hec_heavy_neighbours(int32_t level, int32_t uniform, uint64_t seed, int64_t n, const int64_t* row_offsets,
                             const int32_t* col, const int32_t* ew, const int32_t* vw, int32_t ptr_space,
                             int32_t* hn_out);

/* heuristics.hpp:89-169: aggregate construction from hn[n]. */
int jet_map_construct(int64_t n, const int32_t* hn, int32_t ptr_space, int32_t* map_out, int32_t* nc_out);

/* contract.hpp:347-446: coarse graph for map (n -> nc). out_row_offsets has
 * nc + 1 entries; out_col / out_ew must hold row_offsets[n] entries (an upper
 * bound); rows come out sorted by column. */
int jet_contract(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* vw,
                 const int32_t* map, int32_t nc, int32_t ptr_space, int64_t* out_row_offsets, int32_t* out_col,
                 int32_t* out_ew, int32_t* out_vw, int64_t* out_m);

/* jet_refiner.hpp:958-1048 (init_conn_data): per-vertex part-connectivity
 * tables. out_off[n+1] gives each row's slot range (capacity min(deg, k));
 * out_part/out_val the slots (NULL_PART = -1 when empty); out_size[n] the
 * table sizes. Slot order inside a row is not part of the contract. */
int jet_conn_tables(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                    const int32_t* part, int32_t k, int32_t ptr_space, int64_t* out_off, int32_t* out_part,
                    int32_t* out_val, int32_t* out_size);

/* One move-selection step on fresh refinement state for `part`: kind 0 =
 * jet_lp (jet_refiner.hpp:200-323), 1 = rebalance_weak (:551-645), 2 =
 * rebalance_strong (:417-546). Writes the moves in increasing vertex id and
 * their destinations (n entries each at most). */
int jet_refine_step(int32_t kind, int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                    const int32_t* vw, const int32_t* part, int32_t k, double imb, double filter_ratio,
                    int32_t ptr_space, int32_t* moves_out, int32_t* dest_out, int32_t* nmoves_out);

/* jet_refiner.hpp:1050-1142 on one graph with fresh refine_data. part_io in:
 * start partition, out: best partition. rec_out[8]: 2*cut, total_imb, n, m,
 * iterations, lp_iterations, opt, size_max. */
int jet_refine_level(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* vw,
                     int32_t k, double imb, double tol, int32_t ultra, int32_t uniform, int32_t ptr_space,
                     int32_t* part_io, int64_t* rec_out);

/* jet_refine_level with the iteration controller's record of every iteration
 * (jet_refiner.hpp:1102-1134), for parity tests; no reference counterpart.
 * trace_out holds max_iters records of 6 int64: step kind (0 jet_lp, 1
 * rebalance_weak, 2 rebalance_strong), -1, 2*cut and total_imb after the step,
 * the best-state decision (0 none, 1 better balance, 2 better cut) and the
 * stall counter after it. *iters_out = the number of iterations (records
 * beyond max_iters are dropped). The level runs on the multi-kernel path. */
int jet_refine_level_trace(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                           const int32_t* vw, int32_t k, double imb, double tol, int32_t ultra, int32_t uniform,
                           int32_t ptr_space, int32_t* part_io, int64_t* rec_out, int64_t* trace_out,
                           int32_t max_iters, int32_t* iters_out);

/* uncoarsen.hpp:82-86: fine_out[i] = coarse[map[i]]. */
int jet_project(int64_t n_fine, const int32_t* map, const int32_t* coarse, int32_t ptr_space, int32_t* fine_out);

/* part_stat.hpp:77-108: *cut2_out = sum over directed edges of w*[p_i != p_j]. */
int jet_edge_cut(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* part,
                 int32_t ptr_space, int64_t* cut2_out);

/* jet_edge_cut without any allocation. It has no counterpart in the
 * reference: it serves callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every array is device memory (ptr_space JET_PTR_DEVICE implied; ew may be
 * NULL for unit weights); *cut2_out is host memory. Validation and result are
 * those of jet_edge_cut. The only device memory it uses besides the arguments
 * is a fixed 256-byte array that each GPU holds from the library's first use
 * until the process ends (error word, weight totals, cut accumulator); its
 * size does not depend on n or m, so no graph is too large for it. Calls on
 * one device are serialised by a per-device lock; it calls no hipMalloc,
 * hipFree or allocating routine and runs on the null stream. */
int jet_edge_cut_noalloc(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                         const int32_t* part, int64_t* cut2_out);

/* part_stat.hpp:146-153: sizes_out[k] = vertex weight per part. */
int jet_part_sizes(int64_t n, const int32_t* vw, const int32_t* part, int32_t k, int32_t ptr_space,
                   int64_t* sizes_out);

/* jet_refiner.hpp:326-356 evaluated on the GPU for count (gain, vwgt) pairs
 * (host arrays). */
int jet_gain_bucket_batch(int64_t count, const int64_t* gains, const int64_t* vwgts, int32_t* out);

/* ---- Synthetic graph generators (host memory; the measured workloads). */

/* nx x ny 4-neighbour grid, id y*nx+x, unit weights. rowptr[n+1], col[m]
 * with m = 2*(nx*(ny-1) + ny*(nx-1)). */
int jet_generate_grid2d(int64_t nx, int64_t ny, int64_t* rowptr, int32_t* col);

/* nx x ny x nz 7-point stencil, id (z*ny+y)*nx+x, edge weight
 * 1 + splitmix64(seed ^ (min*N + max)) % wmax (wmax <= 0: unit weights). */
int jet_generate_stencil3d(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, int32_t wmax, int64_t* rowptr,
                           int32_t* col, int32_t* ew);

/* R-MAT (a, b, c, 1-a-b-c) with 2^scale vertices and ef*2^scale drawn edges,
 * self-loops dropped, symmetrised, duplicates removed, rows sorted. col must
 * hold 2*ef*2^scale entries; *m_out receives the stored edge count. */
int jet_generate_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, int64_t* rowptr,
                      int32_t* col, int64_t* m_out);

/* ---- Files and quality metrics of the reference's drivers (app/io.hpp,
 * app/part_eval.cpp, src/part_stat.hpp). */

/* A CSR graph in host memory allocated by the library (jet_free_graph). */
typedef struct jet_host_graph_t {
    int64_t n;
    int64_t m; /* stored entries: both directions */
    int64_t* row_offsets;
    int32_t* col;
    int32_t* ew; /* NULL when the file has no edge weights */
    int32_t uniform_ew;
} jet_host_graph_t;

/* app/io.hpp:96-205: METIS graph file (header "n m [fmt [ncon]]", '%'
 * comments, 1-based neighbours, fmt 1 = edge weights). Vertex weights and
 * ncon are JET_EUNSUPPORTED, as in the reference. */
int jet_load_metis(const char* path, jet_host_graph_t* out);
void jet_free_graph(jet_host_graph_t* g);

/* app/io.hpp:50-81: config file of 4 values (coarsening_alg, num_parts,
 * num_iter, max_imb_ratio) and an optional 5th (ultra_settings); the other
 * fields get config_t's defaults. */
int jet_load_config(const char* path, jet_config_t* cfg);

/* app/io.hpp:207-235: one part id per line. */
int jet_write_partition(const char* path, int64_t n, const int32_t* part);
int jet_read_partition(const char* path, int64_t n, int32_t* part);

/* Coarse-hierarchy dump and replay (src/binary_dump.hpp:61-97,
 * app/import_coarse.cpp:49-145, uncoarsen.hpp:94-111).
 *
 * With config.dump_coarse set, a plain partition call writes two files in the
 * working directory before it refines the first level (from the coarsest)
 * whose partition is balanced: coarse_part.out (that level's partition, raw
 * int32) and coarse_graphs.out (int32 level count, then per level, finest
 * first: n, m, row offsets[n+1], columns[m], edge weights[m], vertex
 * weights[n] and, from the second level on, the map from the previous level).
 * The widths are the reference graph type's for the entry point used:
 * jet_partition_csr32 writes JET_DUMP_32, jet_partition_csr and
 * jet_partition_csr_big JET_DUMP_BIG, jet_partition_csr_biggest
 * JET_DUMP_BIGGEST. Each attempt dumps; the last one's files remain.
 * dump_coarse is rejected for jobs of a jet_ctx. */
#define JET_DUMP_32 0      /* matrix_t: 4-byte ordinals, offsets and weights */
#define JET_DUMP_BIG 1     /* big_matrix_t: 4-byte ordinals, 8-byte offsets and weights */
#define JET_DUMP_BIGGEST 2 /* biggest_matrix_t: 8-byte ordinals, offsets and weights */

typedef struct jet_hierarchy jet_hierarchy_t;

/* import_coarse.cpp:49-110: reads and checks a hierarchy file into host memory. */
int jet_load_hierarchy(const char* path, int32_t kind, jet_hierarchy_t** out);
int32_t jet_hierarchy_levels(const jet_hierarchy_t* h);
int64_t jet_hierarchy_vertices(const jet_hierarchy_t* h, int32_t level); /* level 0 = finest; -1 if out of range */
int64_t jet_hierarchy_edges(const jet_hierarchy_t* h, int32_t level);
void jet_free_hierarchy(jet_hierarchy_t* h);

/* import_coarse.cpp:112-121: n raw int32 part ids. */
int jet_load_coarse_part(const char* path, int64_t n, int32_t* part);

/* import_coarse.cpp:123-145 (uncoarsen.hpp:127-155) on the GPU: refines the
 * loaded hierarchy from its coarsest level, whose partition is coarse_part
 * (host, n of the last level), to the finest; part_out (host) receives the
 * finest level's partition and cut_out its edge cut. Refinement draws no
 * random numbers, so no seed is taken. dump_coarse must be 0. */
int jet_uncoarsen_hierarchy(const jet_config_t* cfg, const jet_hierarchy_t* h, const int32_t* coarse_part,
                            int32_t* part_out, int64_t* cut_out, jet_stats_t* stats);

/* src/part_stat.hpp:77-239 and app/part_eval.cpp:46-115, computed on the GPU. */
typedef struct jet_quality_t {
    int64_t edge_cut;     /* undirected */
    int64_t comm_size;    /* sum over vertices of the other parts among their neighbours */
    int64_t max_part_cut; /* largest cut weight incident to one part */
    int64_t largest_part;
    int64_t smallest_part;
    double max_imb; /* largest / ceil(total / k) */
    double min_imb; /* smallest / ceil(total / k) */
} jet_quality_t;

int jet_partition_quality(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew,
                          const int32_t* vw, const int32_t* part, int32_t k, int32_t ptr_space, jet_quality_t* out);

/* src/part_stat.hpp:116-129 (cut_heatmap), on the GPU: heat_out[p * k + q] =
 * total weight of the edges from vertices of part p to vertices of part q,
 * q != p (each undirected cut edge counted once in each direction; the
 * diagonal is 0). k * k entries, in ptr_space. */
int jet_cut_heatmap(int64_t n, const int64_t* row_offsets, const int32_t* col, const int32_t* ew, const int32_t* part,
                    int32_t k, int32_t ptr_space, int64_t* heat_out);

/* ---- Batched and served partitioning (one GPU, several jobs in flight).
 *
 * A context owns `workers` GPU workers, each with its own HIP stream and
 * device memory, and one host thread per worker. Jobs queue in submission
 * order and start when their estimated device footprint fits the context's
 * budget beside the running jobs, so jobs overlap on the device within a
 * memory bound. Each job runs exactly jet_partition_csr with its own seed on
 * private state, so its result equals the same call made alone. The
 * reference has no batched or served mode: these entry points have no
 * counterpart in it and serve callers that partition many graphs. */
#define JET_MAX_WORKERS 64

typedef struct jet_ctx jet_ctx_t;

/* Seconds since the context was created. */
typedef struct jet_job_times_t {
    double submit_s;
    double start_s;
    double end_s;
} jet_job_times_t;

/* workers in [1, JET_MAX_WORKERS]; budget_bytes <= 0 means 80% of the device's free memory. */
int jet_ctx_create(int32_t device, int32_t workers, int64_t budget_bytes, jet_ctx_t** out);
/* Waits for running jobs, then frees the context (uncollected results are
 * dropped). No jet_submit, jet_poll or jet_partition_batch call on the same
 * context may run concurrently with it or start after it. */
int jet_ctx_destroy(jet_ctx_t* ctx);

/* Queues one job with jet_partition_csr's arguments. The arrays (host or
 * device per ptr_space) and part_out must stay valid until jet_poll returns
 * the job's result. JET_ENOMEM when the job alone exceeds the budget. */
int jet_submit(jet_ctx_t* ctx, const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int32_t* col,
               const int32_t* ew, const int32_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
               int32_t* part_out, int64_t* ticket_out);

/* Result of a job: JET_OK (writes *cut_out, stats, times; the ticket is then
 * consumed), the job's error code, or JET_EAGAIN while it is queued or
 * running. wait_ms: 0 polls, > 0 waits up to that long, < 0 waits until done.
 * One thread at a time may wait for a ticket: a second waiter gets JET_EINVAL. */
int jet_poll(jet_ctx_t* ctx, int64_t ticket, int32_t wait_ms, int64_t* cut_out, jet_stats_t* stats,
             jet_job_times_t* times);

/* Partitions `count` graphs given as graph-tagged concatenated arrays and
 * waits for all of them. Graph g owns vertices [vtx_offsets[g],
 * vtx_offsets[g+1]) of vw / part_out and edges [edge_offsets[g],
 * edge_offsets[g+1]) of col / ew; its n_g + 1 local row offsets (starting at
 * 0) are at row_offsets + vtx_offsets[g] + g; its columns are local ids.
 * cfgs[g], seeds[g] (NULL: 0) and uniform_ew[g] (NULL: 0) are per graph;
 * part_out receives local part ids, cuts_out[g] the cut, stats[g] (NULL: not
 * written) the record. Returns the first failing graph's code. */
int jet_partition_batch(jet_ctx_t* ctx, int32_t count, const jet_config_t* cfgs, const uint64_t* seeds,
                        const int64_t* vtx_offsets, const int64_t* edge_offsets, const int64_t* row_offsets,
                        const int32_t* col, const int32_t* ew, const int32_t* vw, const int32_t* uniform_ew,
                        int32_t ptr_space, int32_t* part_out, int64_t* cuts_out, jet_stats_t* stats);

/* ---- Tuning parameters: launch shapes and path thresholds that change the
 * speed but never the result (every setting gives the same partition). They
 * are process-wide and have no counterpart in the reference. At the first
 * context creation the library applies the tuning table of the device's
 * architecture, tuning/<arch>.txt beside the library (or $JET_TUNING_FILE),
 * written by the autotuner (jet_amd.autotune); without one the built-in
 * defaults stay. Names:
 * "wg_max_n": levels of up to this many vertices are refined by one
 *   workgroup that runs the whole iteration loop (0 = never);
 * "batch_wg_max_n": the same threshold for jobs of jet_partition_batch and
 *   jet_submit on contexts of at least 8 workers, whose one-workgroup levels
 *   are launched together (0 = never, the default);
 * "wg_big_deg": in that kernel, items above this degree take a whole wave;
 * "small_level_n": levels treated as small (wave-built connectivity rows,
 *   wider afterburner lane groups, 128-thread rebalancer select workgroups);
 * "ab_lanes_small", "upd_lanes": lane-group multipliers (1-4) of the
 *   afterburner on small levels and of the update passes;
 * "upd_big_slices": waves (1-64) that share the edges of one move of degree
 *   above 64 in the update passes;
 * "lp_one_pass_n": LP select runs one pass per wave up to this level size;
 * "ct_wave_max", "ct_block_max": contraction rows of up to this many fine
 *   entries take the one-wave (<= 128) or one-workgroup (<= 4096) LDS table;
 * "redraw_lds_max": strong rebalancing puts up to this many (<= 8192)
 *   re-drawn vertices in id order in LDS, longer lists in global memory;
 * "sel_max_blocks": the rebalancers' select kernels run at most this many
 *   workgroups (64-65536), each over a contiguous range of vertices;
 * "sel_copy_wgs": and flush their bucket sums into one copy of the bucket
 *   array per this many workgroups (1-65536; at most 16 copies).
 * jet_tuning_set / jet_tuning_get return JET_EINVAL for an unknown name or a
 * value out of range; jet_tuning_load applies a table file ("name value"
 * lines, '#' comments) and returns JET_EINVAL, applying nothing, when it
 * cannot be read or holds an unknown name or a value out of range. */
int jet_tuning_set(const char* name, int64_t value);
int jet_tuning_get(const char* name, int64_t* value);
int jet_tuning_load(const char* path);

/* ---- Kernel-family timing and algorithmic bytes for measurement tools (no
 * counterpart in the reference). Families: "lp_select", "afterburner",
 * "update", "weak_select"; a unit is one kernel launch of the family.
 * jet_prof_enable("<family>") times every later launch of that family with a
 * pair of HIP events that the launch records at its kernel's start and end
 * (hipExtLaunchKernelGGL), each interval starting no earlier than a marker
 * event recorded just before the launch, so the time is the kernels'
 * execution;
 * "<family>:count" or "all:count" instead counts the launches and the items
 * each family's algorithmic bytes follow from (a separate run: counting adds
 * work); "none" stops. jet_prof_read returns, for an enabled family, its
 * launches and their total milliseconds (timing) or its launches and
 * algorithmic bytes (counting). */
int jet_prof_enable(const char* family);
int jet_prof_read(const char* family, int64_t* units, double* total_ms, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* JET_CAPI_H */
