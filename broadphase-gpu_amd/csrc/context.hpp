// context.hpp -- the library context (one per stq_create): device, stream,
// grow-only scratch, learned-sizing state, batched-pipeline plans and graphs,
// and the autotuning table.  Calls on one context are serialised by the
// caller.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "batch.hpp"
#include "sizer.hpp"
#include "stq_internal.hpp"
#include "sweep.hpp"

namespace stqamd {

// Learned sizing: Adam steps (minibatches of up to 8192 items) and learning
// rate per training scene; slots per item = min(candidates, scale * prediction
// + slack).  One scene: 4 * p + 256.  Any overflowed item costs the exact
// refill, a launch as long as the sweep of the densest such item by one wave,
// while unused slots cost memory only (compaction reads the pairs, not the
// slots): generous slots for a scratch of about four times the pairs.
// Batched scenes: 2 * p + 32; there overflow is rare and the scratch budget of
// the whole batch stays at twice its pairs.
constexpr int kTrainSteps = 64;
constexpr float kTrainRate = 2e-3f;
constexpr float kCapScale = 4.0f;
constexpr float kCapSlack = 256.f;
constexpr float kBatchCapScale = 2.0f;
constexpr float kBatchCapSlack = 32.f;
// Replay ring of training rows: each training scene adds a strided sample of
// up to kReplayPerScene of its items.
constexpr int64_t kReplay = 262144;
constexpr int64_t kReplayPerScene = 8192;

// Planning window when none is set: the FP64 form of batched scenes, and the
// reduced-precision form (whose groups are cut from windows ordered by
// group_key: a larger window gives groups a smaller union box).
constexpr int kBatchWindow = 256;
constexpr int kQWindow = 4096;
// The batched reduced-precision planner's window (and the autotuner's
// starting point there): the count-only tuning trials cannot tell 2048 from
// 4096 (within 2%), while the learned pass on the 512-scene batch takes
// 7.24 ms with 2048 and 7.74 ms with 4096.
constexpr int kBatchQWindow = 2048;
// Reduced-precision item chunk when none is set (positions per body item).
constexpr int kQChunk = 16384;

// One batch shape (scene offsets, window, chunk): its host and device tables
// and the graph captured for it.
struct BatchPlan {
  std::vector<int64_t> offsets;
  int W = 0, chunk = 0;
  BatchTables t;
  BatchDev dev;
  int64_t item_bound = 0;  // no more items than this, whatever the boxes
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_sig = 0;  // launch signature the graph was captured with
  uint64_t eager_sig = 0;  // signature of the last eager run
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_sig = 0;
  }
};

// Grow-only pinned host buffer (results read back after a batched call).
struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t need) {
    if (need > bytes) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      bytes = 0;
      if (hipHostMalloc(&p, need, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        throw StatusError(STQ_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
      }
      bytes = need;
    }
    return p;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Device counters of one call: [0] error flags, [1] candidates, [2] lane
// steps, [4] overflowed items, [15] the mesh calls' error
// flag; from [kClaimBase] the item-claim counters of the first and second
// sweep launch, 8 per launch (one per XCD slice), each alone in a 256-byte
// line (device-scope atomics on one line serialise across the XCDs).
constexpr int kFlagCounter = 15;
constexpr int kClaimBase = 16;
constexpr int kCounters = kClaimBase + 2 * 8 * kClaimStride / 2;
// Per-XCD work claim counters of sweep launch 0 or 1 of one call (zeroed with
// the other counters at its start).
inline unsigned *claim_counters(unsigned long long *ctr, int launch) {
  return reinterpret_cast<unsigned *>(ctr + kClaimBase) + launch * 8 * kClaimStride;
}
// u64 words of the 8 claim lines of launch 0, which also hold the per-XCD
// FP64 confirmation counters (second u64 of each line).
constexpr int kClaimLinesU64 = 8 * kClaimStride / 2;
inline unsigned long long *confirm_counters(unsigned long long *ctr) { return ctr + kClaimBase + 1; }
// Sum of the confirmation counters in a host copy of those lines.
inline unsigned long long confirm_sum(const unsigned long long *lines) {
  unsigned long long s = 0;
  for (int x = 0; x < 8; ++x) s += lines[x * (kClaimStride / 2) + 1];
  return s;
}

enum Ev {
  E_START, E_KEYS, E_SORT, E_GATHER, E_PLAN, E_PRED, E_COUNT, E_SCAN, E_FILL, E_COMPACT, E_END,
  E_NUM
};

// Launch parameters the autotuner picks for one call shape class.
struct TuneEntry {
  int axis_policy, window, chunk, batch, waves_per_cu, tail_cost;
  double ms;  // the winning trial's count-only time
};
// Shape class of a call: precision, batched or not, floor(log2 boxes),
// floor(log2 scenes).
uint64_t tune_key(int precision, bool batched, int64_t n, int nscenes);

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  stq_options opt{};
  stq_stats stats{};
  std::string last_error;
  int cus = 256;
  hipEvent_t ev[E_NUM] = {};
  bool graph_events = false;  // a stream capture can record the phase events
  // The allocation-free entry: this device's fixed scratch array (looked up
  // at stq_create), and whether planning caps the items per group (it does
  // there, so the item count has a bound that the array is sized for).
  void *noalloc_base = nullptr;
  bool plan_cap = false;
  // Sorted rows whose pairs a single-scene call emits: [row_begin, row_end)
  // (row_end < 0: all), the reference's tidstart range.
  int64_t row_begin = 0, row_end = -1;

  DevBuf keys, vals, smin, smax, geo, meta, full, hi, rowlist, groups, nchunks, ioff, items,
      counts, offsets, scan_tmp, counters, axis_partial, h2d_boxes, d2h_pairs, tgroups, gchunks;
  DevBuf item_cost, cost_off, xcd_slices;  // the work split by estimated cost
  SortScratch sort;
  // learned sizing: parameters, Adam state, per-item buffers
  DevBuf sz_params, sz_m, sz_v, sz_grad, sz_partial, sz_loss, feat, cand, pred, caps, cap_off,
      targets, bfeat, btargets, scratch_pairs, ovf, mae_acc, replay_feat, replay_t, col_stats,
      col_work, scene_params, row_feat_pos, row_feat_grp;
  bool sz_ready = false;
  int adam_t = 0;
  int64_t replay_n = 0, replay_pos = 0, train_calls = 0;
  int64_t scenes_seen = 0;
  bool learned_pass = false;  // the last call used the single learned pass
  // batched pipeline
  DevBuf b_tables, b_skey, b_orig, b_pairoff, mesh_boxes;
  DevBuf box_kinds;  // narrow-phase hand-off: one kind byte per box
  PinnedBuf b_host;
  // Small results of a single-scene call (counters, counts, totals) read back
  // into pinned memory: the copies are plain DMA, with no staging through
  // pageable memory on the host.
  PinnedBuf s_host;
  // [0..32) single values, then a copy of the claim lines of launch 0
  int64_t *small_host() {
    return static_cast<int64_t *>(s_host.get((32 + kClaimLinesU64) * sizeof(int64_t)));
  }
  BatchPlan bplan;
  int64_t b_budget = 0;      // scratch pairs of the learned batched pass
  // Most scratch pairs of a learned pass under options.scratch_limit_mb.
  int64_t scratch_limit_pairs() const {
    return opt.scratch_limit_mb > 0 ? std::max<int64_t>(((int64_t)opt.scratch_limit_mb << 20) / 8, 1)
                                    : INT64_MAX;
  }
  int64_t s_last_pairs = 0;  // pairs of the last single-scene call
  int64_t b_last_pairs = 0;  // pairs and items of the last batched call
  int64_t b_last_items = 0;
  // autotuning: launch parameters per call shape class (tune.cpp)
  std::map<uint64_t, TuneEntry> tune_table;
  double last_tune_ms = 0;

  ~Context() {
    bplan.drop_graph();
    b_host.release();
    s_host.release();
    DevBuf *bb[] = {&b_tables, &b_skey,  &b_orig,     &b_pairoff,
                    &tgroups,  &gchunks, &mesh_boxes, &box_kinds};
    for (DevBuf *b : bb) b->release();
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    DevBuf *bufs[] = {&keys,  &vals,    &smin,    &smax,     &geo,          &meta,
                      &full,  &hi,      &rowlist, &groups,   &nchunks,      &ioff,
                      &items, &counts,  &offsets, &scan_tmp, &counters,     &axis_partial,
                      &h2d_boxes, &d2h_pairs, &item_cost, &cost_off, &xcd_slices};
    for (DevBuf *b : bufs) b->release();
    sort.alt_keys.release();
    sort.alt_vals.release();
    sort.counts.release();
    sort.hist.release();
    sort.misc.release();
    DevBuf *szb[] = {&sz_params, &sz_m, &sz_v, &sz_grad, &sz_partial, &sz_loss, &feat, &cand,
                     &pred, &caps, &cap_off, &targets, &bfeat, &btargets, &scratch_pairs, &ovf,
                     &mae_acc, &replay_feat, &replay_t, &col_stats, &col_work, &scene_params,
                     &row_feat_pos, &row_feat_grp};
    for (DevBuf *b : szb) b->release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  void finish_stats(hipStream_t st);
  // Phase times of the last call from the events (learned_pass selects the
  // learned layout of the sweep phases).
  void phase_times();
  void sizer_ensure();
  // Enqueues the exact refill of the overflowed items ovf[0 .. *novf) (at
  // most nbound) of the sweep described by sa: a fill pass over them at their
  // exact offsets off[item] into out (nothing is written when *total >
  // capacity).  No host synchronisation; sweep launch 1.
  void enqueue_refill(const SweepArgs &sa, bool f32, int batch, int nblocks, const int *ovf,
                      const unsigned long long *novf, int64_t nbound, const int64_t *off,
                      int2 *out, const int64_t *total, int64_t capacity, hipStream_t st);
  // XCD slices of equal estimated cost for a sweep launch over sa's items
  // (at most `bound`, or *sa.nitems_dev): the cost of an item counts its
  // shape and its expected pairs, the network's predictions (pred, bounded by
  // the candidates cand) or the exact counts (counts), or neither.  Enqueued
  // only, on buffers that a capture must have reserved
  // (reserve_cost_slices); null when the split is by item count
  // (options.balance = 2 or no XCD split).
  const int64_t *cost_slices(const SweepArgs &sa, int64_t bound, const float *pred,
                             const uint32_t *cand, const uint32_t *counts, hipStream_t st);
  void reserve_cost_slices(int64_t bound);
  void run_learned(SweepArgs sa, bool f32, int batch, int nblocks, const float *d_feat,
                   const uint32_t *d_cand, stq_pair *out, int64_t capacity, stq_pair **alloc_out,
                   int64_t *count);
  void sizer_train(const float *d_feat, const float *d_target, int64_t n, int steps, float lr);
  void sizer_remember(const float *d_feat, const float *d_target, int64_t n, int64_t start = 0);
  float sizer_gradient(const float *d_feat, const float *d_target, int64_t n);
  void run(const stq_aabb_f64 *boxes, int64_t n, stq_pair *out, int64_t capacity,
           stq_pair **alloc_out, int64_t *count);
  // run() with count-then-fill sizing and every scratch buffer carved from the
  // fixed device array (noalloc_base): no device or host memory is allocated.
  void run_noalloc(const stq_aabb_f64 *boxes, int64_t n, stq_pair *out, int64_t capacity,
                   int64_t *count);
  void batch_plan(const int64_t *off, int nscenes, int W, int chunk);
  void run_batched(const stq_aabb_f64 *boxes, const int64_t *off, int nscenes, stq_pair *out,
                   int64_t capacity, int64_t *pair_off);
  // With opt.autotune: tunes the call's shape class on this input if the
  // table has no entry yet (trial runs, results discarded), then copies the
  // entry into opt.  The caller restores opt after the call.
  // trials = false: only an existing table entry is applied.
  void autotune_and_apply(const stq_aabb_f64 *boxes, int64_t n, const int64_t *off, int nscenes,
                          bool batched, bool trials = true);
  int tune_save(const char *path) const;
  int tune_load(const char *path);
};

// Restores a context's options when a call ends (the autotuner overrides
// them for the duration of one call).
struct OptionsScope {
  Context *c;
  stq_options saved;
  explicit OptionsScope(Context *ctx) : c(ctx), saved(ctx->opt) {}
  ~OptionsScope() { c->opt = saved; }
};

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int d) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != d) STQ_HIP(hipSetDevice(d));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace stqamd
