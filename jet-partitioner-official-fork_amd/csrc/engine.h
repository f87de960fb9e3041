// Partition engine shared by the plain entry points (jet_capi.cpp) and the
// batched / served layer (batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "arena.h"
#include "jet_capi.h"

namespace jet {

// One independent GPU worker: a HIP stream (default priority) and two device
// arenas, one for the level hierarchy and one for scratch. Calls on one
// context are serialised by its mutex; different contexts run concurrently.
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena persist;
    Arena scratch;
    std::mutex mu;
};

// Throws on failure (no device, bad index, HIP error).
Context* context_create(int device);
void context_destroy(Context* c);

// jet_partition_csr on the given context. Returns a JET_* code; the message of
// a failure is left for jet_last_error() on the calling thread.
int partition_in_context(Context& C, const jet_config_t* cfg, int64_t n, const int64_t* row_offsets,
                         const int32_t* col, const int32_t* ew, const int32_t* vw, int32_t uniform_ew,
                         int32_t ptr_space, uint64_t seed, int32_t* part_out, int64_t* cut_out, jet_stats_t* stats);

// Sets this thread's jet_last_error() message.
void set_last_error(const std::string& msg);

}  // namespace jet
