// Input staging helpers (device).
#pragma once
#include "arena.h"
#include "jet_types.h"

namespace jet {

// Returns error bits: 1 offsets not monotone, 2 column out of range,
// 4 non-positive edge weight, 8 non-positive vertex weight, 16 weight total
// does not fit int32.
int32_t validate_graph(vid_t n, const eoff_t* rowptr, const vid_t* col, const wgt_t* ew, const wgt_t* vw, Arena& a,
                       hipStream_t s);
// validate_graph's two passes into caller-provided device words (err: one
// int32, zero on entry; sums: two uint64 edge / vertex weight totals, zero on
// entry; NULL weights count as unit). They allocate nothing.
void launch_validate_offsets(vid_t n, const eoff_t* rowptr, eoff_t m, int32_t* err, hipStream_t s);
void launch_validate(vid_t n, eoff_t m, const vid_t* col, const wgt_t* ew, const wgt_t* vw, int32_t* err,
                     unsigned long long* sums, hipStream_t s);
void widen_offsets(int64_t n1, const int32_t* in, int64_t* out, hipStream_t s);
int32_t narrow_weights(int64_t cnt, const int64_t* in, int32_t* out, Arena& a, hipStream_t s);
// int64 columns to int32; returns 2 when a column is outside [0, n).
int32_t narrow_columns(int64_t cnt, const int64_t* in, int64_t n, int32_t* out, Arena& a, hipStream_t s);
// Whether every part id is in [0, k).
bool parts_in_range(int64_t n, const int32_t* part, int32_t k, Arena& a, hipStream_t s);
int64_t conn_capacity(vid_t n, const eoff_t* rowptr, int32_t k, Arena& a, hipStream_t s);
void gain_bucket_batch(int64_t cnt, const int64_t* g, const int64_t* w, int32_t* out, hipStream_t s);

}  // namespace jet
