// sweep.hpp -- record types and launchers of sweep.hip.
#pragma once

#include "stq_internal.hpp"

namespace stqamd {

// One sorted box as the sweep reads it: bounds on the two non-sweep axes
// (the Scalar2 min/max of the reference MiniBox, src/aabb.cuh:75-97), the
// three vertex ids and the type bit.  FP64 form (48 B) and FP32-outward
// form (32 B); the hot loop fetches a whole record with scalar loads.
struct alignas(16) Rec64 {
  double lo1, hi1, lo2, hi2;
  int v0, v1, v2, type;
};
struct alignas(16) Rec32 {
  float lo1, hi1, lo2, hi2;
  int v0, v1, v2, type;
};
// The FP64 box of one sorted position, sweep axis first (FP32-confirm form).
struct alignas(16) Full64 {
  double lo0, hi0, lo1, hi1, lo2, hi2;
};

void launch_axis_stats(const stq_aabb_f64 *boxes, int64_t n, double *partial, int nblocks,
                       hipStream_t st);
void launch_keys(bool f32, const stq_aabb_f64 *boxes, int64_t n, int axis, void *keys,
                 uint32_t *vals, unsigned *err, hipStream_t st);
void launch_gather(bool f32, const stq_aabb_f64 *boxes, const uint32_t *order, int64_t n,
                   int axis, void *smin, void *smax, void *rec, Full64 *full, hipStream_t st);
void launch_chain_end(bool f32, const void *smin, const void *smax, int64_t n, int *hi,
                      unsigned long long *cand, hipStream_t st);
void launch_plan(const int *hi, int64_t n, int W, int chunk, int *rowlist, int4 *groups,
                 uint32_t *nchunks, unsigned long long *lane_steps, hipStream_t st);
void launch_items(const int4 *groups, const int64_t *ioff, int64_t ngroups, int chunk,
                  Item *items, hipStream_t st);
// Count pass (fill = false): counts[it] = pairs of item it.  Fill pass: pairs
// of item it go to out[offsets[it] ..) as (input id of the row, sorted q);
// launch_translate then turns them into canonical input-id pairs.
void launch_sweep(bool f32, bool fill, int nblocks, const Item *items, int64_t nitems,
                  const int *rowlist, const int *hi, const void *rec, const Full64 *full,
                  const uint32_t *orig, uint32_t *counts, const int64_t *offsets, stq_pair *out,
                  unsigned long long *confirms, hipStream_t st);
void launch_translate(stq_pair *pairs, int64_t count, const uint32_t *orig, hipStream_t st);

}  // namespace stqamd
