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
// Reduced-precision form: the two non-sweep axes quantised outward to 15-bit
// levels, packed (axis 1 in the low half, axis 2 in the high half) for the
// lower and the upper bounds; and the exact FP64 box (sweep axis first) with
// its ids for the confirmation of every candidate the levels let through.
struct alignas(8) QRec {
  uint32_t lo, hi;
};
struct alignas(16) Exact64 {
  double lo0, hi0, lo1, hi1, lo2, hi2;
  int v0, v1, v2, type;
};
// Quantisation of one scene: level(x) = (x - origin) * scale per axis;
// scale 0 puts every box on the full level range.
struct QParams {
  double origin[3];
  double scale[3];
};

// Per block: 12 sums (per axis: sum(min+max), sum(min^2+max^2), sum(max-min),
// sum(centre^2)) then per axis min(min) and max(max): kAxisStats doubles.
constexpr int kAxisStats = 18;
// Outward 15-bit levels: t = (x - origin) * scale is monotone in x under IEEE
// rounding, so floor(t(lo)) <= ceil(t(hi')) whenever lo <= hi': a pair that
// overlaps in FP64 always overlaps in levels (a superset test).
__device__ inline uint32_t level_down(double x, double o, double sc) {
  if (sc == 0.0) return 0u;
  const double t = fmin(fmax(floor((x - o) * sc), 0.0), 32767.0);
  return (uint32_t)t;
}
__device__ inline uint32_t level_up(double x, double o, double sc) {
  if (sc == 0.0) return 32767u;
  const double t = fmin(fmax(ceil((x - o) * sc), 0.0), 32767.0);
  return (uint32_t)t;
}

// Per-scene sweep axis and level range (batched scenes).
struct SceneParams {
  int axis;
  int pad;
  QParams qp;
};

// Scene of a sorted or input position: the s with seg_off[s] <= pos <
// seg_off[s+1] (seg_off has nscenes+1 entries, empty scenes are skipped).
__device__ inline int scene_of(const int64_t *__restrict__ seg_off, int nscenes, int64_t pos) {
  int lo = 0, hi = nscenes;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (seg_off[mid] <= pos) lo = mid; else hi = mid;
  }
  return lo;
}

// Chain end: the last q in [s, end) with smin[q] <= x, given smin[s] <= x and
// smin non-decreasing.  Each round issues K independent probes (indices
// clamped into [lo, up), so every load is in bounds) and keeps the bracket
// smin[lo] <= x, up == end or smin[up] > x: galloping with the step growing
// (K+1)-fold, then (K+1)-ary cuts.  A chain of L takes about 2 log_{K+1} L
// dependent load rounds instead of the 2 log_2 L of a binary search.
template <class T>
__device__ __forceinline__ int64_t last_le(const T *__restrict__ smin, int64_t s, int64_t end,
                                           T x) {
  constexpr int K = 4;
  int64_t lo = s, up = end, t = 1;
  while (true) {
    int c = 0;
#pragma unroll
    for (int i = 1; i <= K; ++i) {
      const int64_t p = lo + t * i;
      c += (p < up) & (smin[p < up ? p : up - 1] <= x);
    }
    if (c < K) {
      const int64_t nlo = lo + t * c, nup = lo + t * (c + 1);
      lo = nlo;
      up = nup < up ? nup : up;
      break;
    }
    lo += t * K;
    t *= K + 1;
  }
  while (up - lo > 1) {
    const int64_t st = (up - lo + K) / (K + 1);
    int c = 0;
#pragma unroll
    for (int i = 1; i <= K; ++i) {
      const int64_t p = lo + st * i;
      c += (p < up) & (smin[p < up ? p : up - 1] <= x);
    }
    const int64_t nlo = lo + st * c, nup = lo + st * (c + 1);
    lo = nlo;
    up = nup < up ? nup : up;
  }
  return lo;
}

void launch_axis_stats(const stq_aabb_f64 *boxes, int64_t n, double *partial, int nblocks,
                       hipStream_t st);
void launch_keys(bool f32, const stq_aabb_f64 *boxes, int64_t n, int axis, void *keys,
                 uint32_t *vals, unsigned *err, hipStream_t st);
void launch_gather(bool f32, const stq_aabb_f64 *boxes, const uint32_t *order, int64_t n,
                   int axis, const QParams &qp, void *smin, void *smax, void *rec, QRec *qrec,
                   hipStream_t st);
void launch_chain_end(bool f32, const void *smin, const void *smax, int64_t n, int *hi,
                      unsigned long long *cand, hipStream_t st);
// tail_cost: weight of tail tests in the group cut (plan.hpp); 0 = no tails.
void launch_plan(const int *hi, int64_t n, int W, int chunk, float tail_cost, int *rowlist,
                 int4 *groups, int4 *tgroups, int2 *gchunk, uint32_t *nchunks,
                 unsigned long long *lane_steps, hipStream_t st);
// Items of every group from the group records (body chunks, then tail chunks).
void launch_items(const int4 *groups, const int4 *tgroups, const int2 *gchunk, const int64_t *ioff,
                  int64_t ngroups, Item *items, hipStream_t st);
enum SweepMode { kCount = 0, kFill = 1, kLearned = 2 };

struct SweepArgs {
  const Item *items = nullptr;
  int64_t nitems = 0;            // work items (or selected items when sel != null)
  const int *sel = nullptr;      // optional item selection
  const int *rowlist = nullptr;
  const int *hi = nullptr;
  const void *rec = nullptr;     // Rec64 (FP64 form) or Exact64 by sorted position
  const QRec *qrec = nullptr;    // reduced-precision form only
  const uint32_t *orig = nullptr;
  uint32_t *counts = nullptr;    // kCount, kLearned: exact pairs per item
  const int64_t *offsets = nullptr;  // kFill: output offset; kLearned: scratch offset
  const uint32_t *caps = nullptr;    // kLearned: capacity per item
  int2 *out = nullptr;           // kFill: output; kLearned: scratch
  unsigned long long *confirms = nullptr;
  int zero = 0;                  // always 0 (an opaque runtime zero)
  int xcd_split = 1;             // 1: each XCD sweeps one contiguous slice of items
  // device-resident sizes (batched pipeline): the item count, and for kFill
  // the pair total with the output capacity (nothing is written when the
  // total exceeds it)
  const int64_t *nitems_dev = nullptr;
  const int64_t *total_dev = nullptr;
  int64_t capacity = 0;
  // kXcds zeroed claim counters (one per XCD slice): waves claim items
  // dynamically; null: a fixed stride
  unsigned *work = nullptr;
};

// Pairs come out as canonical (min, max) pairs of input ids (orig[] of both
// sorted positions), in the emission order of the sweep.
void launch_sweep(bool f32, int mode, int batch, int nblocks, const SweepArgs &args,
                  hipStream_t st);

}  // namespace stqamd
