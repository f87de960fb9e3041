// Batched and served partitioning on one GPU.
//
// A jet_ctx owns `workers` engine contexts (one HIP stream and device arena
// each) and one host thread per context. Jobs wait in a FIFO queue and are
// admitted when their estimated HBM footprint fits the context's budget beside
// the jobs already running, so work from several jobs overlaps on the device
// while memory stays bounded. A job runs exactly the single-graph path
// (jet_partition_csr) with its own seed on its worker's private stream and
// arenas: nothing is shared between jobs, so a job's result is the same alone
// and in any mix.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "initpart.h"
#include "refine_wg.h"
#include "jet_capi.h"
#include "jet_types.h"
#include "tuning.h"

namespace {

using clk = std::chrono::steady_clock;

struct Job {
    int64_t ticket = 0;
    jet_config_t cfg{};
    int64_t n = 0;
    const int64_t* rp = nullptr;
    const int32_t* col = nullptr;
    const int32_t* ew = nullptr;
    const int32_t* vw = nullptr;
    int32_t uniform = 0;
    int32_t space = JET_PTR_HOST;
    uint64_t seed = 0;
    int32_t* part_out = nullptr;
    int64_t footprint = 0;
    // a thread is waiting in jet_poll for this job; only that thread may
    // collect (and free) it
    bool claimed = false;
    // results
    bool done = false;
    int rc = JET_EAGAIN;
    std::string err;
    int64_t cut = 0;
    jet_stats_t stats{};
    double t_submit = 0, t_start = 0, t_end = 0;
};

// Gathers the one-workgroup level runs of a context's concurrent jobs into
// graph-tagged launches (k_refine_wg_multi): a run waits until every running
// job has a run pending, the batch is full, or COLLECT_WAIT has passed, and
// whichever waiting job finds the batch ready launches it on the collector's
// own stream for all of them. Each workgroup reads only its own level's
// arrays, so a job's result does not depend on what it was batched with.
class Collector final : public jet::WgCollector {
  public:
    static constexpr int MAX_BATCH = 1024;
    static constexpr auto COLLECT_WAIT = std::chrono::microseconds(300);

    explicit Collector(const std::atomic<int>* running) : running_(running) {}
    ~Collector() override {
        if (d_ps_) (void)hipFree(d_ps_);
        if (d_st_) (void)hipFree(d_st_);
        if (h_ps_) (void)hipHostFree(h_ps_);
        if (h_st_) (void)hipHostFree(h_st_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    void run(const jet::WgParams& p, hipStream_t job_stream, jet::WgState& out) override {
        JET_HIP(hipStreamSynchronize(job_stream));  // the level's tables and state
        Req r{&p, &out};
        std::unique_lock<std::mutex> lk(mu_);
        pending_.push_back(&r);
        cv_.notify_all();
        const auto deadline = clk::now() + COLLECT_WAIT;
        while (!r.done) {
            const bool ready = (int)pending_.size() >= running_->load() || (int)pending_.size() >= MAX_BATCH ||
                               clk::now() >= deadline;
            if (!launching_ && !pending_.empty() && ready) {
                launching_ = true;
                std::vector<Req*> batch;
                batch.swap(pending_);
                lk.unlock();
                std::string err;
                try {
                    launch(batch);
                } catch (const std::exception& e) {
                    err = e.what();
                }
                lk.lock();
                launching_ = false;
                for (Req* q : batch) {
                    q->err = err;
                    q->done = true;
                }
                cv_.notify_all();
            } else if (launching_ || clk::now() >= deadline) {
                cv_.wait(lk);
            } else {
                cv_.wait_until(lk, deadline);
            }
        }
        if (!r.err.empty()) throw jet::DeviceError(r.err);
    }

  private:
    struct Req {
        const jet::WgParams* p;
        jet::WgState* out;
        bool done = false;
        std::string err;
    };

    void launch(const std::vector<Req*>& batch) {
        const int n = (int)batch.size();
        if (!stream_) JET_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        if (n > cap_) {
            if (d_ps_) JET_HIP(hipFree(d_ps_));
            if (d_st_) JET_HIP(hipFree(d_st_));
            if (h_ps_) JET_HIP(hipHostFree(h_ps_));
            if (h_st_) JET_HIP(hipHostFree(h_st_));
            d_ps_ = nullptr;
            d_st_ = nullptr;
            h_ps_ = nullptr;
            h_st_ = nullptr;
            cap_ = 0;
            const int c = std::max(n, 64);
            JET_HIP(hipMalloc(&d_ps_, sizeof(jet::WgParams) * c));
            JET_HIP(hipMalloc(&d_st_, sizeof(jet::WgState) * c));
            JET_HIP(hipHostMalloc(&h_ps_, sizeof(jet::WgParams) * c));
            JET_HIP(hipHostMalloc(&h_st_, sizeof(jet::WgState) * c));
            cap_ = c;
        }
        for (int i = 0; i < n; i++) h_ps_[i] = *batch[i]->p;
        JET_HIP(hipMemcpyAsync(d_ps_, h_ps_, sizeof(jet::WgParams) * n, hipMemcpyHostToDevice, stream_));
        jet::launch_refine_wg_multi(h_ps_, d_ps_, d_st_, n, stream_);
        JET_HIP(hipMemcpyAsync(h_st_, d_st_, sizeof(jet::WgState) * n, hipMemcpyDeviceToHost, stream_));
        JET_HIP(hipStreamSynchronize(stream_));
        for (int i = 0; i < n; i++) *batch[i]->out = h_st_[i];
    }

    const std::atomic<int>* running_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<Req*> pending_;
    bool launching_ = false;
    hipStream_t stream_ = nullptr;
    jet::WgParams* d_ps_ = nullptr;
    jet::WgState* d_st_ = nullptr;
    jet::WgParams* h_ps_ = nullptr;
    jet::WgState* h_st_ = nullptr;
    int cap_ = 0;
};

}  // namespace

struct jet_ctx {
    int device = 0;
    int64_t budget = 0;
    clk::time_point origin;
    std::vector<jet::Context*> engines;
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::unordered_map<int64_t, Job*> jobs;
    int64_t next_ticket = 1;
    int64_t in_use = 0;
    int running = 0;
    std::atomic<int> running_jobs{0};  // `running`, readable without mu
    Collector collector{&running_jobs};
    bool stop = false;

    double now() const { return std::chrono::duration<double>(clk::now() - origin).count(); }
};

namespace {

// Upper estimate of a job's device memory: the staged input (int64 offsets,
// int32 columns and weights), the coarse hierarchy (at most the input again),
// the refiner's tables (8 bytes per slot, slots <= m) and per-vertex arrays,
// and contraction temporaries (radix keys of the spill path).
int64_t footprint_of(int64_t n, int64_t m) { return 128 * n + 48 * m + (int64_t(1) << 20); }

void worker_loop(jet_ctx* X, jet::Context* E) {
    (void)hipSetDevice(X->device);
    // the workers' initial partitions share the host's CPUs
    {
        int cpus = 0;
        cpu_set_t cs;
        if (sched_getaffinity(0, sizeof(cs), &cs) == 0) cpus = CPU_COUNT(&cs);
        if (cpus < 1) cpus = (int)std::thread::hardware_concurrency();
        jet_init::set_thread_budget(std::max(1, cpus / std::max<int>(1, (int)X->engines.size())));
    }
    if ((int)X->engines.size() >= jet::BATCH_WG_MIN_WORKERS) {
        jet::set_batch_thread(true);
        jet::set_wg_collector(&X->collector);
    }
    std::unique_lock<std::mutex> lk(X->mu);
    while (true) {
        // FIFO admission: the head job starts once it fits beside the running
        // jobs (or when nothing runs: a job never waits for an empty device)
        X->cv_work.wait(lk, [&] {
            return X->stop || (!X->queue.empty() &&
                               (X->running == 0 || X->in_use + X->queue.front()->footprint <= X->budget));
        });
        if (X->queue.empty()) {
            if (X->stop) break;
            continue;
        }
        Job* j = X->queue.front();
        X->queue.pop_front();
        X->in_use += j->footprint;
        X->running++;
        X->running_jobs.store(X->running);
        j->t_start = X->now();
        lk.unlock();
        int64_t cut = 0;
        jet_stats_t st;
        int rc = jet::partition_in_context(*E, &j->cfg, j->n, j->rp, j->col, j->ew, j->vw, j->uniform, j->space,
                                           j->seed, j->part_out, &cut, &st);
        std::string err = rc == JET_OK ? std::string() : std::string(jet_last_error());
        lk.lock();
        j->rc = rc;
        j->err = err;
        j->cut = cut;
        j->stats = st;
        j->t_end = X->now();
        j->done = true;
        X->in_use -= j->footprint;
        X->running--;
        X->running_jobs.store(X->running);
        X->cv_done.notify_all();
        X->cv_work.notify_all();
    }
}

// Validates what can be checked on the host without touching the arrays.
int check_job_args(const jet_config_t* cfg, int64_t n, const int64_t* rp, const int32_t* col, int32_t space,
                   int32_t* part_out) {
    if (!cfg || !rp || !col || !part_out || n < 1) {
        jet::set_last_error("config, row_offsets, col and part_out are required and n must be >= 1");
        return JET_EINVAL;
    }
    if (space != JET_PTR_HOST && space != JET_PTR_DEVICE) {
        jet::set_last_error("ptr_space must be JET_PTR_HOST or JET_PTR_DEVICE");
        return JET_EINVAL;
    }
    if (cfg->dump_coarse) {
        // concurrent jobs would write the same two files
        jet::set_last_error("dump_coarse is not supported for jobs of a jet_ctx");
        return JET_EINVAL;
    }
    return JET_OK;
}

int64_t edge_count(const int64_t* rp, int64_t n, int32_t space) {
    if (space == JET_PTR_HOST) return rp[n];
    int64_t m = 0;
    if (hipMemcpy(&m, rp + n, sizeof(m), hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    return m;
}

int submit_locked(jet_ctx* X, Job* j, int64_t* ticket_out) {
    if (j->footprint > X->budget) {
        jet::set_last_error("job footprint exceeds the context's memory budget");
        delete j;
        return JET_ENOMEM;
    }
    j->ticket = X->next_ticket++;
    j->t_submit = X->now();
    X->jobs[j->ticket] = j;
    X->queue.push_back(j);
    *ticket_out = j->ticket;
    X->cv_work.notify_all();
    return JET_OK;
}

}  // namespace

extern "C" {

int jet_ctx_create(int32_t device, int32_t workers, int64_t budget_bytes, jet_ctx_t** out) {
    if (!out) {
        jet::set_last_error("out is NULL");
        return JET_EINVAL;
    }
    *out = nullptr;
    if (workers < 1 || workers > JET_MAX_WORKERS) {
        jet::set_last_error("workers must be in [1, 64]");
        return JET_EINVAL;
    }
    jet_ctx* X = new jet_ctx();
    X->device = device;
    X->origin = clk::now();
    try {
        for (int32_t w = 0; w < workers; w++) X->engines.push_back(jet::context_create(device));
        if (budget_bytes <= 0) {
            size_t fr = 0, tot = 0;
            JET_HIP(hipMemGetInfo(&fr, &tot));
            budget_bytes = (int64_t)(0.8 * (double)fr);
        }
    } catch (const std::exception& e) {
        for (auto* E : X->engines) jet::context_destroy(E);
        delete X;
        jet::set_last_error(e.what());
        return JET_EDEVICE;
    }
    X->budget = budget_bytes;
    for (auto* E : X->engines) X->threads.emplace_back(worker_loop, X, E);
    *out = X;
    return JET_OK;
}

int jet_ctx_destroy(jet_ctx_t* X) {
    if (!X) return JET_OK;
    {
        std::lock_guard<std::mutex> lk(X->mu);
        X->stop = true;
        X->cv_work.notify_all();
    }
    for (auto& t : X->threads) t.join();
    for (auto* E : X->engines) jet::context_destroy(E);
    for (auto& kv : X->jobs) delete kv.second;
    delete X;
    return JET_OK;
}

int jet_submit(jet_ctx_t* X, const jet_config_t* cfg, int64_t n, const int64_t* row_offsets, const int32_t* col,
               const int32_t* ew, const int32_t* vw, int32_t uniform_ew, int32_t ptr_space, uint64_t seed,
               int32_t* part_out, int64_t* ticket_out) {
    if (!X || !ticket_out) {
        jet::set_last_error("context and ticket_out are required");
        return JET_EINVAL;
    }
    int rc = check_job_args(cfg, n, row_offsets, col, ptr_space, part_out);
    if (rc) return rc;
    int64_t m = edge_count(row_offsets, n, ptr_space);
    if (m < 0) {
        jet::set_last_error("could not read row_offsets[n]");
        return JET_EDEVICE;
    }
    Job* j = new Job();
    j->cfg = *cfg;
    j->n = n;
    j->rp = row_offsets;
    j->col = col;
    j->ew = ew;
    j->vw = vw;
    j->uniform = uniform_ew;
    j->space = ptr_space;
    j->seed = seed;
    j->part_out = part_out;
    j->footprint = footprint_of(n, m);
    std::lock_guard<std::mutex> lk(X->mu);
    return submit_locked(X, j, ticket_out);
}

int jet_poll(jet_ctx_t* X, int64_t ticket, int32_t wait_ms, int64_t* cut_out, jet_stats_t* stats,
             jet_job_times_t* times) {
    if (!X) {
        jet::set_last_error("context is NULL");
        return JET_EINVAL;
    }
    std::unique_lock<std::mutex> lk(X->mu);
    auto it = X->jobs.find(ticket);
    if (it == X->jobs.end()) {
        jet::set_last_error("unknown or already collected ticket");
        return JET_EINVAL;
    }
    Job* j = it->second;
    if (j->claimed) {
        jet::set_last_error("another thread is already waiting for this ticket");
        return JET_EINVAL;
    }
    if (!j->done) {
        if (wait_ms == 0) return JET_EAGAIN;
        // while this thread sleeps, the job stays in X->jobs and no other
        // poll may collect it, so `j` stays valid
        j->claimed = true;
        auto pred = [&] { return j->done; };
        bool ok = true;
        if (wait_ms < 0) X->cv_done.wait(lk, pred);
        else ok = X->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
        j->claimed = false;
        if (!ok) return JET_EAGAIN;
    }
    int rc = j->rc;
    if (cut_out) *cut_out = j->cut;
    if (stats) *stats = j->stats;
    if (times) {
        times->submit_s = j->t_submit;
        times->start_s = j->t_start;
        times->end_s = j->t_end;
    }
    if (rc != JET_OK) jet::set_last_error(j->err);
    X->jobs.erase(it);
    delete j;
    return rc;
}

int jet_partition_batch(jet_ctx_t* X, int32_t count, const jet_config_t* cfgs, const uint64_t* seeds,
                        const int64_t* vtx_offsets, const int64_t* edge_offsets, const int64_t* row_offsets,
                        const int32_t* col, const int32_t* ew, const int32_t* vw, const int32_t* uniform_ew,
                        int32_t ptr_space, int32_t* part_out, int64_t* cuts_out, jet_stats_t* stats) {
    if (!X || count < 0 || (count > 0 && (!cfgs || !vtx_offsets || !edge_offsets || !row_offsets || !col ||
                                          !part_out || !cuts_out))) {
        jet::set_last_error("context, configs, offsets, arrays and outputs are required");
        return JET_EINVAL;
    }
    if (ptr_space != JET_PTR_HOST && ptr_space != JET_PTR_DEVICE) {
        jet::set_last_error("ptr_space must be JET_PTR_HOST or JET_PTR_DEVICE");
        return JET_EINVAL;
    }
    std::vector<int64_t> tickets(count, 0);
    int first_rc = JET_OK;
    std::string first_err;
    {
        std::lock_guard<std::mutex> lk(X->mu);
        for (int32_t g = 0; g < count; g++) {
            const int64_t v0 = vtx_offsets[g], n = vtx_offsets[g + 1] - v0;
            const int64_t e0 = edge_offsets[g], m = edge_offsets[g + 1] - e0;
            if (n < 1 || m < 0) {
                first_rc = JET_EINVAL;
                first_err = "graph " + std::to_string(g) + ": empty vertex range or negative edge range";
                break;
            }
            if (cfgs[g].dump_coarse) {
                first_rc = JET_EINVAL;
                first_err = "graph " + std::to_string(g) + ": dump_coarse is not supported for jobs of a jet_ctx";
                break;
            }
            Job* j = new Job();
            j->cfg = cfgs[g];
            j->n = n;
            j->rp = row_offsets + v0 + g;  // graph g's n + 1 local offsets
            j->col = col + e0;
            j->ew = ew ? ew + e0 : nullptr;
            j->vw = vw ? vw + v0 : nullptr;
            j->uniform = uniform_ew ? uniform_ew[g] : 0;
            j->space = ptr_space;
            j->seed = seeds ? seeds[g] : 0;
            j->part_out = part_out + v0;
            j->footprint = footprint_of(n, m);
            int rc = submit_locked(X, j, &tickets[g]);
            if (rc) {
                first_rc = rc;
                first_err = "graph " + std::to_string(g) + ": " + jet_last_error();
                break;
            }
        }
    }
    for (int32_t g = 0; g < count; g++) {
        if (!tickets[g]) continue;
        int rc = jet_poll(X, tickets[g], -1, &cuts_out[g], stats ? &stats[g] : nullptr, nullptr);
        if (rc && first_rc == JET_OK) {
            first_rc = rc;
            first_err = "graph " + std::to_string(g) + ": " + jet_last_error();
        }
    }
    if (first_rc) jet::set_last_error(first_err);
    return first_rc;
}

}  // extern "C"
