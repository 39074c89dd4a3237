// api.cpp -- the C-ABI (include/stq.h): contexts, options and the host
// orchestration of one broad-phase call.
//
// Replaces the host body of the reference's run_sweep_sharedqueue
// (src/simulation.cu:555-782): H2D copy, create_ds, sort, output sizing
// (MemHandler, src/memory.cuh:69-191), twostage_queue, overflow loop.  All
// errors are returned as status codes; scratch is grow-only and owned by the
// context; the output is sized exactly by a count pass instead of from free
// device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "build.hpp"
#include "narrow.hpp"
#include "context.hpp"
#include "noalloc.hpp"
#include "plan.hpp"

namespace stqamd {

static bool options_valid(const stq_options &o) {
  return o.scratch_limit_mb >= 0 && o.balance >= 0 && o.balance <= 2 && o.precision >= 0 &&
         o.precision <= 2 && o.axis_policy >= 0 && o.axis_policy <= 4 &&
         o.sizing >= 0 && o.sizing <= 1 && o.chunk >= 0 && o.window >= 0 &&
         (o.window == 0 || (o.window >= 64 && o.window <= kMaxWindow32 &&
                            (o.window & (o.window - 1)) == 0)) &&
         o.waves_per_cu >= 0 && o.waves_per_cu <= 64 &&
         (o.batch == 0 || o.batch == 2 || o.batch == 4 || o.batch == 8) && o.xcd_split >= 0 &&
         o.xcd_split <= 2 && o.graph >= 0 && o.graph <= 2 && o.claim >= 0 && o.claim <= 2 &&
         o.tails >= 0 && o.tails <= 2 && o.tail_cost >= 0 && o.tail_cost <= 100000;
}

// Weight of a tail test in the group cut; 0 turns tails off.
// Auto (tails = 0) turns tails on for the plain FP64 sweep and off for the
// reduced-precision one, whose body positions mostly cost one union test per
// 64 (a long body is cheap there, a tail item is not).
// The forms whose planner groups rows close on axis 1 (union tests per 64
// positions): reduced precision and grouped FP64.
static bool grouped_form(const stq_options &o) { return o.precision != STQ_PRECISION_F64; }
static float tail_cost(const stq_options &o) {
  const bool on = o.tails == 1 || (o.tails == 0 && !grouped_form(o));
  return !on ? 0.f : (o.tail_cost > 0 ? (float)o.tail_cost * 0.01f : 4.f);
}
// Item chunk when none is set.
static int auto_chunk(const stq_options &o) {
  return o.chunk > 0 ? o.chunk : (grouped_form(o) ? kQChunk : 2048);
}

static float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return ms;
}

void Context::run(const stq_aabb_f64 *boxes, int64_t n, stq_pair *out, int64_t capacity,
                  stq_pair **alloc_out, int64_t *count) {
  stats = stq_stats{};
  stats.n_boxes = n;
  stats.n_scenes = 1;
  stats.precision = opt.precision;
  stats.sizing = opt.sizing;
  const bool f32 = opt.precision == STQ_PRECISION_F32_CONFIRM;
  const bool grouped = grouped_form(opt);
  hipStream_t st = stream;
  STQ_HIP(hipEventRecord(ev[E_START], st));
  if (n <= 1) {
    *count = 0;
    if (alloc_out) *alloc_out = static_cast<stq_pair *>(nullptr);
    for (int e = E_KEYS; e < E_NUM; ++e) STQ_HIP(hipEventRecord(ev[e], st));
    STQ_HIP(hipStreamSynchronize(st));
    stats.axis = opt.axis_policy <= 2 ? opt.axis_policy : 0;
    return;
  }
  // counters: [0] error flags, [1] candidates, [2] lane steps; confirmations
  // per XCD slice in the claim lines
  unsigned long long *ctr = counters.as<unsigned long long>(kCounters);
  STQ_HIP(hipMemsetAsync(ctr, 0, kCounters * sizeof(unsigned long long), st));

  // Axis choice and, for the reduced-precision form and learned sizing, the
  // quantisation range: per-block statistics, combined in block order on the
  // device (k_axis_final), which the keys, the gather and the features read.
  const bool learned = opt.sizing == STQ_SIZING_LEARNED;
  const int pol = opt.axis_policy;
  const int axis = pol <= 2 ? pol : 0;  // where no statistics are made
  SceneParams *d_sp = nullptr;
  uint32_t *order = vals.as<uint32_t>((size_t)n);
  // reduced precision: the statistics pass writes the keys of all three axes
  // and the sort reads those of the chosen axis
  void *kbuf = keys.get((size_t)n * (f32 ? 12 : 8));
  if (pol > 2 || grouped || learned) {
    // about four boxes per thread (the loads of one thread overlap)
    const int nb = (int)std::min<int64_t>((n + 1023) / 1024, 4096);
    double *part = axis_partial.as<double>((size_t)nb * kAxisStats);
    d_sp = scene_params.as<SceneParams>(1);
    launch_axis_stats(boxes, n, part, nb, st, f32 ? static_cast<uint32_t *>(kbuf) : nullptr,
                      (unsigned *)&ctr[0]);
    launch_axis_final(part, nb, n, pol, d_sp, st);
    if (f32) launch_sweep_range(boxes, nullptr, 1, n, d_sp, st);
  }
  stats.axis = axis;
  if (!f32) launch_keys(f32, boxes, n, axis, d_sp, kbuf, order, (unsigned *)&ctr[0], st);
  STQ_HIP(hipEventRecord(ev[E_KEYS], st));
  if (f32) {
    // the sweep keys (sweep_key_bits(n) bits, made from the FP32 keys of the
    // chosen axis in the first pass): every digit a pass, with no digit
    // histogram and no host read-back of it
    const int64_t ntiles = (n + sort_tile_keys() - 1) / sort_tile_keys();
    radix_sort_u32_segmented_axes(static_cast<uint32_t *>(kbuf), order, n, sweep_key_bits(n),
                                  nullptr, ntiles, d_sp, sort, st);
  } else {
    int passes = 0;
    radix_sort_u64((uint64_t *)kbuf, order, n, 64, sort, st, &passes);
  }
  STQ_HIP(hipEventRecord(ev[E_SORT], st));

  const size_t ts = f32 ? 4 : 8;
  void *d_smin = smin.get((size_t)n * ts);
  void *d_smax = smax.get((size_t)n * ts);
  void *d_rec = geo.get((size_t)n * (f32 ? sizeof(Exact64) : sizeof(Rec64)));
  QRec *d_qrec = grouped ? meta.as<QRec>((size_t)n + kQPad) : nullptr;
  int *d_hi = hi.as<int>((size_t)n);
  ColStats *d_cols = learned ? col_stats.as<ColStats>((size_t)(n + kColBlock - 1) / kColBlock)
                             : nullptr;
  // learned sizing: row features by sorted position, then in group order
  uint2 *d_rowf = learned ? row_feat_pos.as<uint2>((size_t)n) : nullptr;
  uint2 *d_rowfeat = learned ? row_feat_grp.as<uint2>((size_t)n) : nullptr;
  launch_gather(f32, boxes, order, n, axis, d_sp, d_smin, d_smax, d_rec, d_qrec, d_cols, d_rowf,
                st);
  launch_chain_end(f32, d_smin, d_smax, n, d_hi, &ctr[1], st, row_begin, row_end);
  STQ_HIP(hipEventRecord(ev[E_GATHER], st));
  // The error flags and the candidate count are read with the item count
  // below, except where the window depends on the mean chain length (the
  // kernels in between stay in bounds on any input).
  int64_t *sh = small_host();  // [0..3] counters, [4] axis, [5] items, [6] total, ...
  const unsigned long long *hc = reinterpret_cast<const unsigned long long *>(sh);
  int *axis_h = reinterpret_cast<int *>(sh + 4);
  auto read_counters = [&]() {
    *axis_h = axis;
    STQ_HIP(hipMemcpyAsync(sh, ctr, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    if (d_sp) STQ_HIP(hipMemcpyAsync(axis_h, &d_sp->axis, sizeof(int), hipMemcpyDeviceToHost, st));
    STQ_HIP(hipStreamSynchronize(st));
    stats.axis = *axis_h;
    if (hc[0]) throw StatusError(STQ_ERR_INVALID_BOX, "input box with NaN or min > max");
    stats.n_candidates = (int64_t)hc[1];
  };

  // plan
  int W = opt.window;
  if (W == 0 && grouped) W = kQWindow;
  const bool early = W == 0;
  if (early) {
    read_counters();
    const double mean = (double)hc[1] / (double)n;
    W = 256;
    while (W < 2048 && (double)(2 * W) <= mean / 8.0) W *= 2;
  }
  if (!grouped) W = std::min(W, kMaxWindow64);  // 64-bit window keys
  const int chunk = auto_chunk(opt);
  const int64_t ngroups = (n + 63) / 64;
  int2 *d_rowlist = rowlist.as<int2>((size_t)n);
  int4 *d_groups = groups.as<int4>((size_t)ngroups);
  int4 *d_tg = tgroups.as<int4>((size_t)ngroups);
  int2 *d_gch = gchunks.as<int2>((size_t)ngroups);
  uint32_t *d_nch = nchunks.as<uint32_t>((size_t)ngroups);
  int64_t *d_ioff = ioff.as<int64_t>((size_t)ngroups + 1);
  launch_plan(d_hi, d_qrec, n, W, chunk, tail_cost(opt), d_rowlist, d_groups, d_tg, d_gch, d_nch,
              &ctr[2], st, plan_cap, d_rowf, d_rowfeat);
  scan_u32_to_i64(d_nch, d_ioff, ngroups, scan_tmp, st);
  STQ_HIP(hipMemcpyAsync(sh + 5, d_ioff + ngroups, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  if (early)
    STQ_HIP(hipStreamSynchronize(st));
  else
    read_counters();
  const int64_t nitems = sh[5];
  stats.n_items = nitems;
  Item *d_items = items.as<Item>((size_t)std::max<int64_t>(nitems, 1));
  launch_items(d_groups, d_tg, d_gch, d_ioff, ngroups, d_items, st);
  STQ_HIP(hipEventRecord(ev[E_PLAN], st));

  const int wpc = opt.waves_per_cu > 0 ? opt.waves_per_cu : 32;
  // a multiple of 8 blocks so every XCD gets the same number
  const int nblocks = (int)(8 * std::max<int64_t>(
      1, std::min<int64_t>((nitems + 31) / 32, (int64_t)cus * std::max(1, wpc / 4) / 8)));
  uint32_t *d_counts = counts.as<uint32_t>((size_t)std::max<int64_t>(nitems, 1));
  int64_t *d_off = offsets.as<int64_t>((size_t)nitems + 1);
  SweepArgs sa;
  sa.items = d_items;
  sa.nitems = nitems;
  sa.rowlist = d_rowlist;
  sa.rowfeat = d_rowfeat;
  sa.rec = d_rec;
  sa.qrec = d_qrec;
  sa.orig = order;
  sa.counts = d_counts;
  sa.confirms = confirm_counters(ctr);
  sa.xcd_split = opt.xcd_split == 2 ? 0 : 1;
  const bool claim = opt.claim != 2;
  sa.work = claim ? claim_counters(ctr, 0) : nullptr;
  const int batch = opt.batch ? opt.batch : (f32 ? 8 : 2);

  // learned sizing: features of every item; the network predicts, or (while
  // training) learns from the exact counts of this call
  float *d_feat = nullptr;
  uint32_t *d_cand = nullptr;
  const int train_scenes = opt.train_scenes > 0 ? opt.train_scenes : (opt.train_scenes < 0 ? 0 : 200);
  const bool use_net = learned && nitems > 0;
  learned_pass = use_net && scenes_seen >= train_scenes;
  stats.n_tiles = use_net ? nitems : 0;
  if (use_net) {
    sizer_ensure();
    d_feat = feat.as<float>((size_t)nitems * kFeat);
    d_cand = cand.as<uint32_t>((size_t)nitems);
    const int64_t ncols = (n + kColBlock - 1) / kColBlock;
    launch_features(f32, sa, d_cols, ncols, col_work.as<double>(col_work_doubles(ncols)),
                    QParams{}, 0, d_sp, nullptr, 1, d_feat, d_cand, st);
    ++scenes_seen;
  }
  if (learned_pass) {
    run_learned(sa, f32, batch, nblocks, d_feat, d_cand, out, capacity, alloc_out, count);
    return;
  }
  STQ_HIP(hipEventRecord(ev[E_PRED], st));
  sa.slices = cost_slices(sa, nitems, nullptr, nullptr, nullptr, st);
  launch_sweep(f32, kCount, batch, nblocks, sa, st);
  STQ_HIP(hipEventRecord(ev[E_COUNT], st));
  scan_u32_to_i64(d_counts, d_off, nitems, scan_tmp, st);
  STQ_HIP(hipMemcpyAsync(sh + 6, d_off + nitems, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  STQ_HIP(hipEventRecord(ev[E_SCAN], st));
  STQ_HIP(hipStreamSynchronize(st));
  const int64_t total = sh[6];
  *count = total;
  stats.n_pairs = total;
  s_last_pairs = total;
  if (use_net) {
    // prediction error before this scene's training, then the training steps
    float *d_pred = pred.as<float>((size_t)nitems);
    const float *w = sz_params.as<float>(kParams);
    launch_mlp_forward(w, d_feat, nitems, d_pred, st);
    double *acc = mae_acc.as<double>(1);
    STQ_HIP(hipMemsetAsync(acc, 0, sizeof(double), st));
    launch_abs_error(d_pred, d_counts, nitems, acc, st);
    float *d_t = targets.as<float>((size_t)nitems);
    launch_targets(d_counts, nitems, d_t, st);
    // a sample of this scene joins the replay ring; training draws from the
    // whole ring, so earlier scenes are not forgotten
    sizer_remember(d_feat, d_t, nitems);
    sizer_train(replay_feat.as<float>((size_t)kReplay * kFeat), replay_t.as<float>(kReplay), replay_n,
                kTrainSteps, kTrainRate);
    STQ_HIP(hipMemcpyAsync(sh + 8, acc, sizeof(double), hipMemcpyDeviceToHost, st));
    STQ_HIP(hipStreamSynchronize(st));
    stats.sizing_mae = *reinterpret_cast<const double *>(sh + 8) / (double)nitems;
  }
  stq_pair *dst = out;
  if (alloc_out) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(stq_pair) * (size_t)std::max<int64_t>(total, 1)) != hipSuccess) {
      (void)hipGetLastError();
      throw StatusError(STQ_ERR_OUT_OF_MEMORY, "output allocation failed");
    }
    dst = static_cast<stq_pair *>(p);
    *alloc_out = dst;
  } else if (total > capacity) {
    for (int e = E_FILL; e < E_NUM; ++e) STQ_HIP(hipEventRecord(ev[e], st));
    finish_stats(st);
    throw StatusError(STQ_ERR_CAPACITY, "output capacity too small");
  }
  if (total > 0) {
    sa.slices = cost_slices(sa, nitems, nullptr, nullptr, d_counts, st);
    sa.counts = nullptr;
    sa.confirms = nullptr;
    sa.offsets = d_off;
    sa.out = reinterpret_cast<int2 *>(dst);
    sa.work = claim ? claim_counters(ctr, 1) : nullptr;
    launch_sweep(f32, kFill, batch, nblocks, sa, st);
  }
  STQ_HIP(hipEventRecord(ev[E_FILL], st));
  STQ_HIP(hipEventRecord(ev[E_COMPACT], st));
  STQ_HIP(hipEventRecord(ev[E_END], st));
  finish_stats(st);
}

void Context::run_noalloc(const stq_aabb_f64 *boxes, int64_t n, stq_pair *out, int64_t capacity,
                          int64_t *count) {
  if (!noalloc_base) throw StatusError(STQ_ERR_INTERNAL, "no fixed scratch on this device");
  Arena arena;
  arena.base = static_cast<char *>(noalloc_base);
  arena.size = kNoAllocScratchBytes;
  // every buffer a count-then-fill call of run() touches, swapped for an
  // arena-backed one for this call and put back afterwards
  DevBuf *bufs[] = {&keys,     &vals,          &smin,           &smax,          &geo,
                    &meta,     &hi,            &rowlist,        &groups,        &tgroups,
                    &gchunks,  &nchunks,       &ioff,           &items,         &counts,
                    &offsets,  &scan_tmp,      &counters,       &axis_partial,  &scene_params,
                    &sort.alt_keys, &sort.alt_vals, &sort.counts, &sort.hist,    &sort.misc,
                    &item_cost, &cost_off, &xcd_slices};
  constexpr int kBufs = sizeof(bufs) / sizeof(bufs[0]);
  struct Swap {
    DevBuf **b;
    DevBuf saved[kBufs];
    Swap(DevBuf **bb, Arena *a) : b(bb) {
      for (int i = 0; i < kBufs; ++i) {
        saved[i] = *b[i];
        *b[i] = DevBuf{};
        b[i]->arena = a;
      }
    }
    ~Swap() {
      for (int i = 0; i < kBufs; ++i) *b[i] = saved[i];
    }
  } swap(bufs, &arena);
  struct Cap {
    bool &f;
    explicit Cap(bool &x) : f(x) { f = true; }
    ~Cap() { f = false; }
  } cap(plan_cap);
  run(boxes, n, out, capacity, nullptr, count);
}

void Context::reserve_cost_slices(int64_t bound) {
  item_cost.as<uint32_t>((size_t)std::max<int64_t>(bound, 1));
  cost_off.as<int64_t>((size_t)bound + 1);
  xcd_slices.as<int64_t>(kXcds + 1);
  scan_reserve_i64(bound, scan_tmp);
}

const int64_t *Context::cost_slices(const SweepArgs &sa, int64_t bound, const float *pred,
                                    const uint32_t *cand, const uint32_t *counts, hipStream_t st) {
  // auto: only before the learned pass (measured: the network's estimate
  // moves 14% off the batched sweep; before a count or fill pass the shape
  // and exact counts alone misjudge the per-position cost of clustered scenes)
  if (!sa.xcd_split || opt.balance == 2 || bound <= 0 || (opt.balance == 0 && !pred))
    return nullptr;
  reserve_cost_slices(bound);
  uint32_t *c = item_cost.as<uint32_t>((size_t)bound);
  int64_t *pre = cost_off.as<int64_t>((size_t)bound + 1);
  int64_t *sl = xcd_slices.as<int64_t>(kXcds + 1);
  launch_item_cost(sa.items, bound, sa.nitems_dev, sa.qrec != nullptr, pred, cand, counts, c, st);
  scan_u32_to_i64(c, pre, bound, scan_tmp, st, sa.nitems_dev);
  launch_xcd_slices(pre, bound, sa.nitems_dev, sl, st);
  return sl;
}

void Context::enqueue_refill(const SweepArgs &sa, bool f32, int batch, int nblocks, const int *ovf,
                             const unsigned long long *novf, int64_t nbound, const int64_t *off,
                             int2 *out, const int64_t *total, int64_t capacity, hipStream_t st) {
  SweepArgs fa = sa;
  fa.sel = ovf;
  fa.slices = nullptr;
  fa.nitems = nbound;
  fa.nitems_dev = reinterpret_cast<const int64_t *>(novf);
  fa.counts = nullptr;
  fa.confirms = nullptr;
  fa.caps = nullptr;
  fa.offsets = off;
  fa.out = out;
  fa.total_dev = total;
  fa.capacity = capacity;
  if (fa.work) fa.work = claim_counters(counters.as<unsigned long long>(kCounters), 1);
  launch_sweep(f32, kFill, batch, nblocks, fa, st);
}

// The single emitting pass sized by the network.  Every item gets
// min(candidates, kCapScale * predicted + kCapSlack) slots of a scratch
// buffer; the pass
// writes what fits and counts everything, so the exact counts (and the
// output offsets) come out of the same pass.  Items whose count exceeded
// their slots are redone by the exact fill pass straight into the output;
// the rest are copied.  The output order is fixed by the exact counts alone,
// so it is the same as the count-then-fill path's.
void Context::run_learned(SweepArgs sa, bool f32, int batch, int nblocks, const float *d_feat,
                          const uint32_t *d_cand, stq_pair *out, int64_t capacity,
                          stq_pair **alloc_out, int64_t *count) {
  hipStream_t st = stream;
  const int64_t nitems = sa.nitems;
  const float *w = sz_params.as<float>(kParams);
  float *d_pred = pred.as<float>((size_t)nitems);
  uint32_t *d_caps = caps.as<uint32_t>((size_t)nitems);
  int64_t *d_capoff = cap_off.as<int64_t>((size_t)nitems + 1);
  launch_mlp_forward(w, d_feat, nitems, d_pred, st);
  launch_caps(d_pred, d_cand, nitems, kCapScale, kCapSlack, d_caps, st);
  scan_u32_to_i64(d_caps, d_capoff, nitems, scan_tmp, st);
  // The scratch buffer holds exactly the predicted slots (their total is read
  // back once), so an item is redone by the exact refill only when it has
  // more pairs than its prediction allows -- or, with options.scratch_limit_mb,
  // when its slots end beyond that limit.
  int64_t *sh = small_host();
  STQ_HIP(hipMemcpyAsync(sh + 9, d_capoff + nitems, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  STQ_HIP(hipStreamSynchronize(st));
  const int64_t slots = sh[9];
  const int64_t budget = std::min<int64_t>(std::max<int64_t>(slots, 1), scratch_limit_pairs());
  if (budget < slots) launch_cap_budget(d_caps, d_capoff, nitems, budget, st);
  int2 *d_scratch = scratch_pairs.as<int2>((size_t)budget);
  sa.slices = cost_slices(sa, nitems, d_pred, d_cand, nullptr, st);
  STQ_HIP(hipEventRecord(ev[E_PRED], st));
  STQ_HIP(hipEventRecord(ev[E_COUNT], st));
  uint32_t *d_counts = sa.counts;
  int64_t *d_off = offsets.as<int64_t>((size_t)nitems + 1);
  sa.offsets = d_capoff;
  sa.caps = d_caps;
  sa.out = d_scratch;
  launch_sweep(f32, kLearned, batch, nblocks, sa, st);
  STQ_HIP(hipEventRecord(ev[E_FILL], st));
  scan_u32_to_i64(d_counts, d_off, nitems, scan_tmp, st);
  const int64_t *d_total = d_off + nitems;
  double *acc = mae_acc.as<double>(1);
  STQ_HIP(hipMemsetAsync(acc, 0, sizeof(double), st));
  launch_abs_error(d_pred, d_counts, nitems, acc, st);
  STQ_HIP(hipEventRecord(ev[E_SCAN], st));
  // Compaction and the exact refill check the device total against the
  // capacity themselves (nothing is written when it does not fit), and the
  // refill reads the number of overflowed items from the device: the only
  // host synchronisation left is the final one.  A library-allocated output
  // needs the total first.
  stq_pair *dst = out;
  int64_t cap_dev = capacity;
  if (alloc_out) {
    STQ_HIP(hipMemcpyAsync(sh + 6, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    STQ_HIP(hipStreamSynchronize(st));
    const int64_t total = sh[6];
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(stq_pair) * (size_t)std::max<int64_t>(total, 1)) != hipSuccess) {
      (void)hipGetLastError();
      throw StatusError(STQ_ERR_OUT_OF_MEMORY, "output allocation failed");
    }
    dst = static_cast<stq_pair *>(p);
    *alloc_out = dst;
    cap_dev = total;
  }
  int *d_ovf = ovf.as<int>((size_t)nitems);
  unsigned long long *d_novf = counters.as<unsigned long long>(kCounters) + 4;
  STQ_HIP(hipMemsetAsync(d_novf, 0, sizeof(unsigned long long), st));
  launch_compact(d_counts, d_caps, d_capoff, d_off, nitems, d_scratch,
                 reinterpret_cast<int2 *>(dst), d_ovf, d_novf, st, nullptr, d_total, cap_dev);
  enqueue_refill(sa, f32, batch, nblocks, d_ovf, d_novf, nitems, d_off,
                 reinterpret_cast<int2 *>(dst), d_total, cap_dev, st);
  STQ_HIP(hipMemcpyAsync(sh + 6, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  STQ_HIP(hipMemcpyAsync(sh + 7, d_novf, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  STQ_HIP(hipEventRecord(ev[E_COMPACT], st));
  STQ_HIP(hipEventRecord(ev[E_END], st));
  STQ_HIP(hipMemcpyAsync(sh + 8, acc, sizeof(double), hipMemcpyDeviceToHost, st));
  finish_stats(st);
  const int64_t total = sh[6];
  *count = total;
  stats.n_pairs = total;
  s_last_pairs = total;
  stats.n_fallback_tiles = sh[7];
  stats.sizing_mae =
      *reinterpret_cast<const double *>(sh + 8) / (double)std::max<int64_t>(nitems, 1);
  if (!alloc_out && total > capacity) throw StatusError(STQ_ERR_CAPACITY, "output capacity too small");
}

void Context::sizer_ensure() {
  if (sz_ready) return;
  std::vector<float> h(kParams);
  sizer_init_params(h.data(), 0x5717u);
  float *w = sz_params.as<float>(kParams);
  STQ_HIP(hipMemcpyAsync(w, h.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, stream));
  STQ_HIP(hipMemsetAsync(sz_m.as<float>(kParams), 0, sizeof(float) * kParams, stream));
  STQ_HIP(hipMemsetAsync(sz_v.as<float>(kParams), 0, sizeof(float) * kParams, stream));
  STQ_HIP(hipStreamSynchronize(stream));
  adam_t = 0;
  sz_ready = true;
}

float Context::sizer_gradient(const float *d_feat, const float *d_target, int64_t n) {
  float *part = sz_partial.as<float>(mlp_partial_floats(n));
  float *g = sz_grad.as<float>(kParams);
  float *loss = sz_loss.as<float>(1);
  launch_mlp_gradient(sz_params.as<float>(kParams), d_feat, d_target, n, part, g, loss, stream);
  return 0.f;
}

// Adds rows (start + k) * stride mod n, k < kReplayPerScene, to the ring.
void Context::sizer_remember(const float *d_feat, const float *d_target, int64_t n,
                             int64_t start) {
  if (n <= 0) return;
  float *rf = replay_feat.as<float>((size_t)kReplay * kFeat);
  float *rt = replay_t.as<float>((size_t)kReplay);
  const int64_t take = std::min(n, kReplayPerScene);
  int64_t stride = 104729;
  while (std::__gcd(stride, n) != 1) ++stride;
  const int64_t first = std::min(take, kReplay - replay_pos);
  launch_gather_batch(d_feat, d_target, n, start, stride, first, rf + replay_pos * kFeat,
                      rt + replay_pos, stream);
  if (take > first)
    launch_gather_batch(d_feat, d_target, n, start + first, stride, take - first, rf, rt, stream);
  replay_pos = (replay_pos + take) % kReplay;
  replay_n = std::min(kReplay, replay_n + take);
}

// `steps` Adam steps on minibatches of up to 8192 rows taken with a stride
// co-prime to n (a fixed pseudo-random order continued across calls).
void Context::sizer_train(const float *d_feat, const float *d_target, int64_t n, int steps,
                          float lr) {
  if (n <= 0) return;
  const int64_t bs = std::min<int64_t>(n, 8192);
  float *part = sz_partial.as<float>(mlp_partial_floats(bs));
  int64_t stride = 7919;
  while (std::__gcd(stride, n) != 1) ++stride;
  const int64_t call = train_calls++;
  for (int s = 0; s < steps; ++s) {
    // the gradient kernel gathers the minibatch rows itself
    const int64_t start = ((call * steps + s) * bs) % n;
    ++adam_t;
    launch_train_step(sz_params.as<float>(kParams), sz_m.as<float>(kParams), sz_v.as<float>(kParams),
                      d_feat, d_target, n, start, stride, bs, part, lr, adam_t, stream);
  }
}

void Context::finish_stats(hipStream_t st) {
  unsigned long long *hc = reinterpret_cast<unsigned long long *>(small_host() + 10);
  unsigned long long *ctr = counters.as<unsigned long long>(kCounters);
  unsigned long long *lines = reinterpret_cast<unsigned long long *>(small_host() + 32);
  STQ_HIP(hipMemcpyAsync(hc, ctr, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  STQ_HIP(hipMemcpyAsync(lines, ctr + kClaimBase, kClaimLinesU64 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, st));
  STQ_HIP(hipStreamSynchronize(st));
  stats.n_lane_steps = (int64_t)hc[2];
  stats.n_confirm = (int64_t)confirm_sum(lines);
  stats.ms_total = ev_ms(ev[E_START], ev[E_END]);
  phase_times();
}

void Context::phase_times() {
  stats.ms_keys = ev_ms(ev[E_START], ev[E_KEYS]);
  stats.ms_sort = ev_ms(ev[E_KEYS], ev[E_SORT]);
  stats.ms_gather = ev_ms(ev[E_SORT], ev[E_GATHER]);
  stats.ms_plan = ev_ms(ev[E_GATHER], ev[E_PLAN]);
  stats.ms_predict = ev_ms(ev[E_PLAN], ev[E_PRED]);
  if (learned_pass) {
    // predict | (scratch) | learned pass | scan | compact + refill
    stats.ms_sweep_count = 0;
    stats.ms_sweep_fill = ev_ms(ev[E_COUNT], ev[E_FILL]);
    stats.ms_scan = ev_ms(ev[E_FILL], ev[E_SCAN]);
    stats.ms_compact = ev_ms(ev[E_SCAN], ev[E_COMPACT]);
  } else {
    stats.ms_sweep_count = ev_ms(ev[E_PRED], ev[E_COUNT]);
    stats.ms_scan = ev_ms(ev[E_COUNT], ev[E_SCAN]);
    stats.ms_sweep_fill = ev_ms(ev[E_SCAN], ev[E_FILL]);
    stats.ms_compact = 0;
  }
}

// ---- batched scenes ----------------------------------------------------------
// Host tables of a batch shape, uploaded once; the graph of the previous
// shape is dropped when the shape changes.
void Context::batch_plan(const int64_t *off, int nscenes, int W, int chunk) {
  BatchPlan &bp = bplan;
  if (bp.W == W && bp.chunk == chunk && (int)bp.offsets.size() == nscenes + 1 &&
      std::equal(bp.offsets.begin(), bp.offsets.end(), off))
    return;
  bp.drop_graph();
  bp.eager_sig = 0;
  bp.offsets.clear();
  build_batch_tables(off, nscenes, W, bp.t);
  const BatchTables &t = bp.t;
  int64_t bound = 0;
  for (int s = 0; s < nscenes; ++s) {
    const int64_t m = off[s + 1] - off[s];
    if (m < 2) continue;
    const int64_t g = t.scene_first_group[s + 1] - t.scene_first_group[s];
    // body and tail: each at most kCapChunks items of at least `chunk` positions
    bound += 2 * g * std::min<int64_t>(kCapChunks, (m - 1 + chunk - 1) / chunk);
  }
  bp.item_bound = bound;

  const size_t nc = t.chunk_start.size(), nw = t.win_start.size(), ns1 = (size_t)nscenes + 1,
               nst = t.sort_tiles.size();
  size_t at = 0;
  auto place = [&](size_t bytes) {
    const size_t o = at;
    at += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_seg = place(8 * ns1), o_cs = place(8 * nc), o_ce = place(8 * nc),
               o_scb = place(4 * ns1), o_ws = place(8 * nw), o_wc = place(4 * nw),
               o_wg = place(8 * nw), o_sfg = place(8 * ns1),
               o_st = place(sizeof(SegTile) * std::max<size_t>(nst, 1)), upload = at;
  const size_t o_par = place(sizeof(SceneParams) * (size_t)nscenes),
               o_part = place(sizeof(double) * kAxisStats * std::max<size_t>(nc, 1));
  std::vector<char> h(upload);
  auto put = [&](size_t o, const void *src, size_t bytes) {
    if (bytes) std::memcpy(h.data() + o, src, bytes);
  };
  put(o_seg, t.seg_off.data(), 8 * ns1);
  put(o_cs, t.chunk_start.data(), 8 * nc);
  put(o_ce, t.chunk_end.data(), 8 * nc);
  put(o_scb, t.scene_chunk_begin.data(), 4 * ns1);
  put(o_ws, t.win_start.data(), 8 * nw);
  put(o_wc, t.win_cnt.data(), 4 * nw);
  put(o_wg, t.win_gbase.data(), 8 * nw);
  put(o_sfg, t.scene_first_group.data(), 8 * ns1);
  put(o_st, t.sort_tiles.data(), sizeof(SegTile) * nst);
  char *d = static_cast<char *>(b_tables.get(at));
  STQ_HIP(hipMemcpyAsync(d, h.data(), upload, hipMemcpyHostToDevice, stream));
  STQ_HIP(hipStreamSynchronize(stream));
  BatchDev &dv = bp.dev;
  dv.seg_off = reinterpret_cast<int64_t *>(d + o_seg);
  dv.chunk_start = reinterpret_cast<int64_t *>(d + o_cs);
  dv.chunk_end = reinterpret_cast<int64_t *>(d + o_ce);
  dv.scene_chunk_begin = reinterpret_cast<int *>(d + o_scb);
  dv.win_start = reinterpret_cast<int64_t *>(d + o_ws);
  dv.win_cnt = reinterpret_cast<int *>(d + o_wc);
  dv.win_gbase = reinterpret_cast<int64_t *>(d + o_wg);
  dv.scene_first_group = reinterpret_cast<int64_t *>(d + o_sfg);
  dv.sort_tiles = reinterpret_cast<SegTile *>(d + o_st);
  dv.params = reinterpret_cast<SceneParams *>(d + o_par);
  dv.partial = reinterpret_cast<double *>(d + o_part);
  bp.offsets.assign(off, off + nscenes + 1);
  bp.W = W;
  bp.chunk = chunk;
}

static uint64_t fnv1a(const void *p, size_t n, uint64_t h = 1469598103934665603ull) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

// All scenes in one sequence of launches with no host synchronisation until
// the end: per-scene statistics, segmented keys and sort, scene-bounded chain
// ends and planning, then either one count sweep and one fill sweep over
// every scene's items, or (learned sizing) one emitting sweep into predicted
// slots, compaction and the exact refill of overflowed items.  Item and pair
// totals stay on the device; the output is written only when it fits.  The
// sequence is captured in a hipGraph and replayed while the batch shape, the
// buffers and the options are unchanged.
void Context::run_batched(const stq_aabb_f64 *boxes, const int64_t *off_in, int nscenes,
                          stq_pair *out, int64_t capacity, int64_t *pair_off) {
  stats = stq_stats{};
  stats.n_scenes = nscenes;
  stats.precision = opt.precision;
  stats.sizing = opt.sizing;
  stats.axis = opt.axis_policy <= 2 ? opt.axis_policy : -1;
  const bool f32 = opt.precision == STQ_PRECISION_F32_CONFIRM;
  const bool grouped = grouped_form(opt);
  hipStream_t st = stream;
  std::vector<int64_t> off((size_t)nscenes + 1);
  for (int s = 0; s <= nscenes; ++s) off[s] = off_in[s] - off_in[0];
  boxes += off_in[0];
  const int64_t n = off[nscenes];
  if (n > INT32_MAX) throw StatusError(STQ_ERR_TOO_LARGE, "more than 2^31-1 boxes in a batch");
  stats.n_boxes = n;
  const int W = opt.window ? (grouped ? opt.window : std::min(opt.window, kMaxWindow64))
                           : (grouped ? kBatchQWindow : kBatchWindow);
  const int chunk = auto_chunk(opt);
  batch_plan(off.data(), nscenes, W, chunk);
  BatchPlan &bp = bplan;
  const BatchTables &t = bp.t;
  const int64_t ngroups = t.ngroups, bound = bp.item_bound;
  if (bound == 0) {  // no scene holds two boxes
    for (int s = 0; s <= nscenes; ++s) pair_off[s] = 0;
    STQ_HIP(hipEventRecord(ev[E_START], st));
    STQ_HIP(hipEventRecord(ev[E_END], st));
    STQ_HIP(hipStreamSynchronize(st));
    return;
  }
  int sbits = 0;
  while ((1ll << sbits) < nscenes) ++sbits;

  // sizing: count-then-fill, count-then-fill while the network trains on the
  // exact counts, or the single learned pass
  const bool learned = opt.sizing == STQ_SIZING_LEARNED;
  const int train_scenes = opt.train_scenes > 0 ? opt.train_scenes : (opt.train_scenes < 0 ? 0 : 200);
  const bool training = learned && scenes_seen < train_scenes;
  const bool single_pass = learned && !training;
  if (learned) sizer_ensure();
  // the first budget: the slots the previous batch's exact counts would ask
  // for, and a quarter more, since the predictions may ask for a little more
  // and a budget that grows makes a new buffer (the call after it is not
  // captured)
  if (single_pass && b_budget == 0) {
    const int64_t slots = (int64_t)kBatchCapScale * b_last_pairs +
                          (int64_t)kBatchCapSlack * b_last_items;
    b_budget = std::max<int64_t>(int64_t(1) << 22, slots + slots / 4);
  }
  if (single_pass) b_budget = std::min(b_budget, scratch_limit_pairs());

  // every buffer is reserved here: the launches below allocate nothing
  const size_t ts = f32 ? 4 : 8;
  // reduced precision: the keys of all three axes (the sort reads its scene's)
  void *d_keys = keys.get((size_t)n * (f32 ? 12 : 8));
  uint32_t *d_vals = vals.as<uint32_t>((size_t)n);
  uint32_t *d_skey = f32 ? nullptr : b_skey.as<uint32_t>((size_t)n);
  if (f32)
    radix_sort_segmented_reserve(n, (int64_t)t.sort_tiles.size(), sort);
  else
    radix_sort_reserve(n, 8, sort);
  void *d_smin = smin.get((size_t)n * ts);
  void *d_smax = smax.get((size_t)n * ts);
  void *d_rec = geo.get((size_t)n * (f32 ? sizeof(Exact64) : sizeof(Rec64)));
  QRec *d_qrec = grouped ? meta.as<QRec>((size_t)n + kQPad) : nullptr;
  uint32_t *d_orig = b_orig.as<uint32_t>((size_t)n);
  int *d_hi = hi.as<int>((size_t)n);
  int2 *d_rowlist = rowlist.as<int2>((size_t)n);
  int4 *d_groups = groups.as<int4>((size_t)ngroups);
  int2 *d_gchunk = gchunks.as<int2>((size_t)ngroups);
  int4 *d_tg = tgroups.as<int4>((size_t)ngroups);
  uint32_t *d_nch = nchunks.as<uint32_t>((size_t)ngroups);
  int64_t *d_ioff = ioff.as<int64_t>((size_t)ngroups + 1);
  Item *d_items = items.as<Item>((size_t)bound);
  uint32_t *d_counts = counts.as<uint32_t>((size_t)bound);
  int64_t *d_off = offsets.as<int64_t>((size_t)bound + 1);
  scan_reserve_i64(std::max(ngroups, bound), scan_tmp);
  int64_t *d_pairoff = b_pairoff.as<int64_t>((size_t)nscenes + 1);
  unsigned long long *ctr = counters.as<unsigned long long>(kCounters);
  float *d_feat = learned ? feat.as<float>((size_t)bound * kFeat) : nullptr;
  const int64_t ncols = (n + kColBlock - 1) / kColBlock;
  ColStats *d_cols = learned ? col_stats.as<ColStats>((size_t)ncols) : nullptr;
  uint2 *d_rowf = learned ? row_feat_pos.as<uint2>((size_t)n) : nullptr;
  uint2 *d_rowfeat = learned ? row_feat_grp.as<uint2>((size_t)n) : nullptr;
  double *d_colwork = learned ? col_work.as<double>(col_work_doubles(ncols)) : nullptr;
  uint32_t *d_cand = learned ? cand.as<uint32_t>((size_t)bound) : nullptr;
  float *d_pred = learned ? pred.as<float>((size_t)bound) : nullptr;
  uint32_t *d_caps = single_pass ? caps.as<uint32_t>((size_t)bound) : nullptr;
  int64_t *d_capoff = single_pass ? cap_off.as<int64_t>((size_t)bound + 1) : nullptr;
  int *d_ovf = single_pass ? ovf.as<int>((size_t)bound) : nullptr;
  int2 *d_scratch = single_pass ? scratch_pairs.as<int2>((size_t)b_budget) : nullptr;
  double *d_acc = single_pass ? mae_acc.as<double>(1) : nullptr;
  const float *d_w = learned ? sz_params.as<float>(kParams) : nullptr;
  // pinned results: pair offsets, 5 counters, items, total slots, error sum,
  // the claim lines of the first sweep launch (confirmation counters)
  reserve_cost_slices(bound);
  const int64_t nh = (int64_t)nscenes + 1 + 8 + kClaimLinesU64;
  int64_t *hbuf = static_cast<int64_t *>(b_host.get(sizeof(int64_t) * (size_t)nh));
  int64_t *h_ctr = hbuf + nscenes + 1;

  const int wpc = opt.waves_per_cu > 0 ? opt.waves_per_cu : 32;
  const int nblocks = (int)(8 * std::max<int64_t>(
      1, std::min<int64_t>((bound + 31) / 32, (int64_t)cus * std::max(1, wpc / 4) / 8)));
  const int batch = opt.batch ? opt.batch : (f32 ? 8 : 2);

  // Phase events: plain records when eager; inside a capture, event record
  // nodes of the graph (where the runtime supports them), so a replay times
  // its phases too.
  bool capturing = false;
  auto mark = [&](int e) {
    if (!capturing)
      STQ_HIP(hipEventRecord(ev[e], st));
    else if (graph_events)
      STQ_HIP(hipEventRecordWithFlags(ev[e], st, hipEventRecordExternal));
  };
  auto enqueue = [&]() {
    launch_fill_u64(ctr, 0ull, kCounters, st);
    launch_seg_stats(boxes, t, bp.dev, opt.axis_policy, opt.axis_policy <= 2 ? opt.axis_policy : 0,
                     f32 ? static_cast<uint32_t *>(d_keys) : nullptr,
                     reinterpret_cast<unsigned *>(&ctr[0]), st);
    if (!f32)
      launch_keys_seg(f32, boxes, t, bp.dev, d_keys, d_vals, reinterpret_cast<unsigned *>(&ctr[0]),
                      st);
    mark(E_KEYS);
    if (f32) {
      // sweep keys sorted within each scene (the scene order is kept), each
      // scene's made from the three-axis FP32 keys by its parameters
      radix_sort_u32_segmented_axes(static_cast<uint32_t *>(d_keys), d_vals, n,
                                    sweep_key_bits(t.max_scene), bp.dev.sort_tiles,
                                    (int64_t)t.sort_tiles.size(), bp.dev.params, sort, st);
    } else {
      radix_sort_u64_sync_free(static_cast<uint64_t *>(d_keys), d_vals, n, 64, sort, st);
      if (sbits > 0) {
        launch_scene_keys(d_vals, t, bp.dev, d_skey, st);
        radix_sort_u32_sync_free(d_skey, d_vals, n, sbits, sort, st);
      }
    }
    mark(E_SORT);
    launch_gather_seg(f32, boxes, d_vals, t, bp.dev, d_smin, d_smax, d_rec, d_qrec, d_orig, d_cols,
                      d_rowf, st);
    launch_chain_end_seg(f32, d_smin, d_smax, t, bp.dev, d_hi, &ctr[1], st);
    mark(E_GATHER);
    launch_plan_seg(d_hi, d_qrec, t, bp.dev, W, chunk, tail_cost(opt), d_rowlist, d_groups, d_tg,
                    d_gchunk, d_nch, &ctr[2], st, d_rowf, d_rowfeat);
    scan_u32_to_i64(d_nch, d_ioff, ngroups, scan_tmp, st);
    launch_items(d_groups, d_tg, d_gchunk, d_ioff, ngroups, d_items, st);
    mark(E_PLAN);
    // item-level arrays are used up to the device item count only: every item
    // below it gets its count from the sweep, and the scans stop there
    SweepArgs sa;
    sa.items = d_items;
    sa.nitems = bound;
    sa.nitems_dev = d_ioff + ngroups;
    sa.rowlist = d_rowlist;
    sa.rowfeat = d_rowfeat;
    sa.rec = d_rec;
    sa.qrec = d_qrec;
    sa.orig = d_orig;
    sa.counts = d_counts;
    sa.confirms = confirm_counters(ctr);
    sa.xcd_split = opt.xcd_split == 2 ? 0 : 1;
    sa.work = opt.claim != 2 ? claim_counters(ctr, 0) : nullptr;
    if (learned)
      launch_features(f32, sa, d_cols, ncols, d_colwork, QParams{}, 0, bp.dev.params,
                      bp.dev.seg_off, nscenes, d_feat, d_cand, st);
    const int64_t *total = d_off + bound;
    if (!single_pass) {
      mark(E_PRED);
      sa.slices = cost_slices(sa, bound, nullptr, nullptr, nullptr, st);
      launch_sweep(f32, kCount, batch, nblocks, sa, st);
      mark(E_COUNT);
      scan_u32_to_i64(d_counts, d_off, bound, scan_tmp, st, sa.nitems_dev);
      mark(E_SCAN);
      sa.slices = cost_slices(sa, bound, nullptr, nullptr, d_counts, st);
      sa.counts = nullptr;
      sa.confirms = nullptr;
      sa.offsets = d_off;
      sa.out = reinterpret_cast<int2 *>(out);
      sa.total_dev = total;
      sa.capacity = capacity;
      if (sa.work) sa.work = claim_counters(ctr, 1);
      launch_sweep(f32, kFill, batch, nblocks, sa, st);
      mark(E_FILL);
      mark(E_COMPACT);
    } else {
      launch_mlp_forward(d_w, d_feat, bound, d_pred, st, sa.nitems_dev);
      launch_caps(d_pred, d_cand, bound, kBatchCapScale, kBatchCapSlack, d_caps, st, sa.nitems_dev);
      scan_u32_to_i64(d_caps, d_capoff, bound, scan_tmp, st, sa.nitems_dev);
      launch_cap_budget(d_caps, d_capoff, bound, b_budget, st, sa.nitems_dev);
      sa.slices = cost_slices(sa, bound, d_pred, d_cand, nullptr, st);
      mark(E_PRED);
      mark(E_COUNT);
      sa.offsets = d_capoff;
      sa.caps = d_caps;
      sa.out = d_scratch;
      launch_sweep(f32, kLearned, batch, nblocks, sa, st);
      mark(E_FILL);
      scan_u32_to_i64(d_counts, d_off, bound, scan_tmp, st, sa.nitems_dev);
      launch_fill_u64(reinterpret_cast<unsigned long long *>(d_acc), 0ull, 1, st);
      launch_abs_error(d_pred, d_counts, bound, d_acc, st, sa.nitems_dev);
      mark(E_SCAN);
      launch_compact(d_counts, d_caps, d_capoff, d_off, bound, d_scratch,
                     reinterpret_cast<int2 *>(out), d_ovf, &ctr[4], st, sa.nitems_dev, total,
                     capacity);
      // the overflowed items, exactly, straight into the output
      enqueue_refill(sa, f32, batch, nblocks, d_ovf, &ctr[4], bound, d_off,
                     reinterpret_cast<int2 *>(out), total, capacity, st);
      mark(E_COMPACT);
    }
    launch_scene_pairs(d_off, d_ioff, t, bp.dev, d_pairoff, st);
    STQ_HIP(hipMemcpyAsync(hbuf, d_pairoff, sizeof(int64_t) * ((size_t)nscenes + 1),
                           hipMemcpyDeviceToHost, st));
    STQ_HIP(hipMemcpyAsync(h_ctr, ctr, 5 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    STQ_HIP(hipMemcpyAsync(h_ctr + 5, d_ioff + ngroups, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    STQ_HIP(hipMemcpyAsync(h_ctr + 8, ctr + kClaimBase, kClaimLinesU64 * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, st));
    if (single_pass) {
      STQ_HIP(hipMemcpyAsync(h_ctr + 6, d_capoff + bound, sizeof(int64_t), hipMemcpyDeviceToHost, st));
      STQ_HIP(hipMemcpyAsync(h_ctr + 7, d_acc, sizeof(double), hipMemcpyDeviceToHost, st));
    }
  };

  struct {
    const void *boxes, *out, *host;
    int64_t capacity, budget;
    uint64_t generation;
    int32_t precision, axis_policy, batch, nblocks, xcd_split, sizing, claim, tails, balance;
    float tail_cost;
  } key;
  std::memset(&key, 0, sizeof(key));
  key.boxes = boxes;
  key.out = out;
  key.host = hbuf;
  key.capacity = capacity;
  key.budget = single_pass ? b_budget : 0;
  key.generation = g_devbuf_generation.load();
  key.precision = opt.precision;
  key.axis_policy = opt.axis_policy;
  key.batch = batch;
  key.nblocks = nblocks;
  key.xcd_split = opt.xcd_split;
  key.sizing = single_pass ? 1 : 0;
  key.claim = opt.claim;
  key.tails = opt.tails;
  key.balance = opt.balance == 2 ? 2 : 0;
  key.tail_cost = tail_cost(opt);
  const uint64_t sig = fnv1a(&key, sizeof(key)) | 1ull;

  bool done = false;
  if (opt.graph != 2 && !training) {
    if (bp.exec && bp.graph_sig == sig) {
      STQ_HIP(hipEventRecord(ev[E_START], st));
      STQ_HIP(hipGraphLaunch(bp.exec, st));
      stats.graph = 2;
      done = true;
    } else if (opt.graph == 1 || bp.eager_sig == sig) {
      bp.drop_graph();
      hipGraph_t g = nullptr;
      STQ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      capturing = true;
      try {
        enqueue();
      } catch (...) {
        capturing = false;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      capturing = false;
      STQ_HIP(hipStreamEndCapture(st, &g));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        throw HipError(e, "hipGraphInstantiate", __FILE__, __LINE__);
      }
      if (g_devbuf_generation.load() != key.generation) {  // cannot happen: all reserved above
        (void)hipGraphExecDestroy(ex);
        (void)hipGraphDestroy(g);
        throw StatusError(STQ_ERR_INTERNAL, "buffer allocated during graph capture");
      }
      bp.graph = g;
      bp.exec = ex;
      bp.graph_sig = sig;
      STQ_HIP(hipEventRecord(ev[E_START], st));
      STQ_HIP(hipGraphLaunch(ex, st));
      stats.graph = 1;
      done = true;
    }
  }
  if (!done) {
    STQ_HIP(hipEventRecord(ev[E_START], st));
    enqueue();
    bp.eager_sig = sig;
  }
  STQ_HIP(hipEventRecord(ev[E_END], st));
  STQ_HIP(hipStreamSynchronize(st));
  stats.ms_total = ev_ms(ev[E_START], ev[E_END]);
  learned_pass = single_pass;
  // the phase events ran unless a graph without event nodes was launched
  if (stats.graph == 0 || graph_events) phase_times();
  const uint64_t *hc = reinterpret_cast<const uint64_t *>(h_ctr);
  for (int s = 0; s <= nscenes; ++s) pair_off[s] = hbuf[s];
  const int64_t nitems = h_ctr[5], total = pair_off[nscenes];
  stats.n_candidates = (int64_t)hc[1];
  stats.n_lane_steps = (int64_t)hc[2];
  stats.n_confirm = (int64_t)confirm_sum(reinterpret_cast<const unsigned long long *>(h_ctr + 8));
  stats.n_items = nitems;
  stats.n_pairs = total;
  if (hc[0]) throw StatusError(STQ_ERR_INVALID_BOX, "input box with NaN or min > max");
  b_last_pairs = total;
  b_last_items = nitems;
  if (learned) {
    stats.n_tiles = nitems;
    scenes_seen += nscenes;
  }
  if (single_pass) {
    stats.n_fallback_tiles = (int64_t)hc[4];
    double mae;
    std::memcpy(&mae, &h_ctr[7], sizeof(double));
    stats.sizing_mae = mae / (double)std::max<int64_t>(nitems, 1);
    // predicted slots beyond the scratch were redone exactly; the next call
    // gets room for all of them (a new buffer, hence a new graph)
    const int64_t want = h_ctr[6];
    if (want > b_budget) b_budget = std::min(want + want / 8, scratch_limit_pairs());
  }
  if (training && nitems > 0) {
    // error of the prediction before training, then a few training rounds
    // on samples of this batch's items
    launch_mlp_forward(d_w, d_feat, nitems, d_pred, st);
    double *acc = mae_acc.as<double>(1);
    STQ_HIP(hipMemsetAsync(acc, 0, sizeof(double), st));
    launch_abs_error(d_pred, d_counts, nitems, acc, st);
    float *d_t = targets.as<float>((size_t)nitems);
    launch_targets(d_counts, nitems, d_t, st);
    const int rounds = std::max(1, std::min(nscenes, 16));
    for (int r = 0; r < rounds; ++r) {
      sizer_remember(d_feat, d_t, nitems, (int64_t)r * kReplayPerScene);
      sizer_train(replay_feat.as<float>((size_t)kReplay * kFeat), replay_t.as<float>(kReplay),
                  replay_n, kTrainSteps, kTrainRate);
    }
    double mae = 0;
    STQ_HIP(hipMemcpyAsync(&mae, acc, sizeof(double), hipMemcpyDeviceToHost, st));
    STQ_HIP(hipStreamSynchronize(st));
    stats.sizing_mae = mae / (double)nitems;
  }
  if (total > capacity) throw StatusError(STQ_ERR_CAPACITY, "output capacity too small");
}

template <class F>
static int guarded(Context *c, F &&f) {
  if (!c) return STQ_ERR_INVALID_ARGUMENT;
  try {
    DeviceGuard g(c->device);
    f();
    c->last_error.clear();
    return STQ_OK;
  } catch (const StatusError &e) {
    c->last_error = e.what;
    return e.status;
  } catch (const HipError &e) {
    c->last_error = e.what;
    (void)hipGetLastError();
    return STQ_ERR_HIP;
  } catch (const std::bad_alloc &) {
    c->last_error = "host allocation failed";
    return STQ_ERR_OUT_OF_MEMORY;
  } catch (...) {
    c->last_error = "internal error";
    return STQ_ERR_INTERNAL;
  }
}

// One call with the autotuned options (when enabled), restored afterwards;
// the tuning time is reported in the call's statistics.
template <class F>
static void tuned_call(Context *c, const stq_aabb_f64 *boxes, int64_t n, const int64_t *off,
                       int nscenes, bool batched, F &&f) {
  OptionsScope scope(c);
  c->autotune_and_apply(boxes, n, off, nscenes, batched);
  try {
    f();
  } catch (...) {
    c->stats.ms_tune = c->last_tune_ms;
    throw;
  }
  c->stats.ms_tune = c->last_tune_ms;
}

}  // namespace stqamd

using stqamd::Context;

extern "C" {

int stq_options_default(stq_options *opt) {
  if (!opt) return STQ_ERR_INVALID_ARGUMENT;
  std::memset(opt, 0, sizeof(*opt));
  opt->precision = STQ_PRECISION_F64;
  opt->axis_policy = STQ_AXIS_X;
  opt->sizing = STQ_SIZING_COUNT_FILL;
  return STQ_OK;
}

int stq_create(void **ctx, int device, const stq_options *opt) {
  if (!ctx) return STQ_ERR_INVALID_ARGUMENT;
  *ctx = nullptr;
  stq_options o;
  stq_options_default(&o);
  if (opt) o = *opt;
  if (!stqamd::options_valid(o)) return STQ_ERR_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    return ndev == 0 ? STQ_ERR_HIP : STQ_ERR_INVALID_ARGUMENT;
  }
  Context *c = new (std::nothrow) Context();
  if (!c) return STQ_ERR_OUT_OF_MEMORY;
  c->device = device;
  c->opt = o;
  int rc = stqamd::guarded(c, [&] {
    STQ_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) STQ_HIP(hipEventCreate(&e));
    c->cus = stqamd::device_cu_count(device);
    // Can a capture hold event record nodes?  Capture only (nothing runs).
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      const hipError_t e1 = hipEventRecordWithFlags(c->ev[stqamd::E_KEYS], c->stream,
                                                    hipEventRecordExternal);
      const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
      c->graph_events = e1 == hipSuccess && e2 == hipSuccess && g != nullptr;
      if (g) (void)hipGraphDestroy(g);
    }
    (void)hipGetLastError();
    // set-up of the allocation-free entry: this device's scratch address and
    // the pinned buffer of the small read-backs, so that entry allocates
    // nothing later
    const int st = stqamd::noalloc_scratch(device, &c->noalloc_base);
    if (st != STQ_OK) throw stqamd::StatusError(st, "hipGetSymbolAddress of the fixed scratch failed");
    c->small_host();
  });
  if (rc != STQ_OK) {
    delete c;
    return rc;
  }
  *ctx = c;
  return STQ_OK;
}

int stq_destroy(void *ctx) {
  Context *c = static_cast<Context *>(ctx);
  if (!c) return STQ_ERR_INVALID_ARGUMENT;
  {
    stqamd::DeviceGuard g(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
  }
  return STQ_OK;
}

int stq_set_options(void *ctx, const stq_options *opt) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !opt || !stqamd::options_valid(*opt)) return STQ_ERR_INVALID_ARGUMENT;
  c->opt = *opt;
  // a captured batch holds launch arguments of the old options
  c->bplan.drop_graph();
  c->bplan.eager_sig = 0;
  return STQ_OK;
}

int stq_find_pairs_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n, stq_pair *d_out,
                          int64_t capacity, int64_t *count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !count || n < 0 || capacity < 0 || (n > 0 && !d_boxes) || (capacity > 0 && !d_out))
    return STQ_ERR_INVALID_ARGUMENT;
  if (n > INT32_MAX) return STQ_ERR_TOO_LARGE;
  return stqamd::guarded(c, [&] {
    stqamd::tuned_call(c, d_boxes, n, nullptr, 1, false,
                       [&] { c->run(d_boxes, n, d_out, capacity, nullptr, count); });
  });
}

int stq_find_pairs_device_noalloc(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n,
                                  stq_pair *d_out, int64_t capacity, int64_t *count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !count || n < 0 || capacity < 0 || (n > 0 && !d_boxes) || (capacity > 0 && !d_out))
    return STQ_ERR_INVALID_ARGUMENT;
  if (n > stqamd::kNoAllocMaxBoxes) {
    c->last_error = "more boxes than the allocation-free entry's maximum (STQ_NOALLOC_MAX_BOXES)";
    return STQ_ERR_TOO_LARGE;
  }
  return stqamd::guarded(c, [&] {
    stqamd::OptionsScope scope(c);
    c->autotune_and_apply(d_boxes, n, nullptr, 1, false, false);  // a table entry, no trials
    c->opt.sizing = STQ_SIZING_COUNT_FILL;
    c->run_noalloc(d_boxes, n, d_out, capacity, count);
  });
}

int stq_find_pairs_rows_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n,
                               int64_t row_begin, int64_t row_end, stq_pair *d_out,
                               int64_t capacity, int64_t *count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !count || n < 0 || capacity < 0 || (n > 0 && !d_boxes) || (capacity > 0 && !d_out) ||
      row_begin < 0 || row_end < row_begin)
    return STQ_ERR_INVALID_ARGUMENT;
  if (n > INT32_MAX) return STQ_ERR_TOO_LARGE;
  return stqamd::guarded(c, [&] {
    stqamd::OptionsScope scope(c);
    c->opt.precision = STQ_PRECISION_F64;
    c->opt.axis_policy = STQ_AXIS_X;
    c->opt.sizing = STQ_SIZING_COUNT_FILL;
    c->opt.autotune = 0;
    struct Rows {
      Context *c;
      Rows(Context *x, int64_t b, int64_t e) : c(x) {
        c->row_begin = b;
        c->row_end = e;
      }
      ~Rows() {
        c->row_begin = 0;
        c->row_end = -1;
      }
    } rows(c, std::min(row_begin, n), std::min(row_end, n));
    c->run(d_boxes, n, d_out, capacity, nullptr, count);
  });
}

int stq_sort_order_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n, int32_t *d_order) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || n < 0 || (n > 0 && (!d_boxes || !d_order))) return STQ_ERR_INVALID_ARGUMENT;
  if (n > INT32_MAX) return STQ_ERR_TOO_LARGE;
  return stqamd::guarded(c, [&] {
    if (n == 0) return;
    unsigned long long *ctr = c->counters.as<unsigned long long>(stqamd::kCounters);
    STQ_HIP(hipMemsetAsync(ctr, 0, sizeof(unsigned long long), c->stream));
    uint32_t *order = c->vals.as<uint32_t>((size_t)n);
    uint64_t *keys = c->keys.as<uint64_t>((size_t)n);
    stqamd::launch_keys(false, d_boxes, n, 0, nullptr, keys, order, reinterpret_cast<unsigned *>(ctr),
                        c->stream);
    int passes = 0;
    stqamd::radix_sort_u64(keys, order, n, 64, c->sort, c->stream, &passes);
    STQ_HIP(hipMemcpyAsync(d_order, order, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice,
                           c->stream));
    unsigned long long err = 0;
    STQ_HIP(hipMemcpyAsync(&err, ctr, sizeof(err), hipMemcpyDeviceToHost, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
    if (err) throw stqamd::StatusError(STQ_ERR_INVALID_BOX, "input box with NaN or min > max");
  });
}

int stq_radix_sort_device(void *ctx, void *d_keys, uint32_t *d_vals, int64_t n, int32_t key_bytes,
                          int32_t key_bits, int32_t form) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || n < 0 || (n > 0 && (!d_keys || !d_vals)) || (key_bytes != 4 && key_bytes != 8) ||
      key_bits < 1 || key_bits > 8 * key_bytes || form < 0 || form > 2 ||
      (form == 2 && key_bytes != 4))
    return STQ_ERR_INVALID_ARGUMENT;
  if (n > INT32_MAX) return STQ_ERR_TOO_LARGE;
  return stqamd::guarded(c, [&] {
    int passes = 0;
    if (key_bytes == 8) {
      uint64_t *k = static_cast<uint64_t *>(d_keys);
      if (form == 0) stqamd::radix_sort_u64(k, d_vals, n, key_bits, c->sort, c->stream, &passes);
      else stqamd::radix_sort_u64_sync_free(k, d_vals, n, key_bits, c->sort, c->stream);
    } else {
      uint32_t *k = static_cast<uint32_t *>(d_keys);
      if (form == 0) stqamd::radix_sort_u32(k, d_vals, n, key_bits, c->sort, c->stream, &passes);
      else if (form == 1) stqamd::radix_sort_u32_sync_free(k, d_vals, n, key_bits, c->sort, c->stream);
      else stqamd::radix_sort_u32_all_digits(k, d_vals, n, key_bits, c->sort, c->stream);
    }
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_device_memory(void *ctx, size_t *free_bytes, size_t *total_bytes) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !free_bytes || !total_bytes) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] { STQ_HIP(hipMemGetInfo(free_bytes, total_bytes)); });
}

int stq_find_pairs_device_alloc(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n,
                                stq_pair **d_out, int64_t *count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !count || !d_out || n < 0 || (n > 0 && !d_boxes)) return STQ_ERR_INVALID_ARGUMENT;
  if (n > INT32_MAX) return STQ_ERR_TOO_LARGE;
  *d_out = nullptr;
  return stqamd::guarded(c, [&] {
    stqamd::tuned_call(c, d_boxes, n, nullptr, 1, false,
                       [&] { c->run(d_boxes, n, nullptr, 0, d_out, count); });
    if (!*d_out) {
      void *p = nullptr;
      STQ_HIP(hipMalloc(&p, sizeof(stq_pair)));
      *d_out = static_cast<stq_pair *>(p);
    }
  });
}

int stq_find_pairs_host(void *ctx, const stq_aabb_f64 *h_boxes, int64_t n, stq_pair **h_out,
                        int64_t *count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !count || !h_out || n < 0 || (n > 0 && !h_boxes)) return STQ_ERR_INVALID_ARGUMENT;
  if (n > INT32_MAX) return STQ_ERR_TOO_LARGE;
  *h_out = nullptr;
  return stqamd::guarded(c, [&] {
    stq_aabb_f64 *db = c->h2d_boxes.as<stq_aabb_f64>((size_t)std::max<int64_t>(n, 1));
    if (n > 0)
      STQ_HIP(hipMemcpyAsync(db, h_boxes, sizeof(stq_aabb_f64) * n, hipMemcpyHostToDevice,
                             c->stream));
    // first pass sizes the output exactly, so a retry never happens
    int64_t cnt = 0;
    stq_pair *dout = nullptr;
    stqamd::tuned_call(c, db, n, nullptr, 1, false, [&] { c->run(db, n, nullptr, 0, &dout, &cnt); });
    stq_pair *h = static_cast<stq_pair *>(std::malloc(sizeof(stq_pair) * (size_t)std::max<int64_t>(cnt, 1)));
    if (!h) {
      if (dout) (void)hipFree(dout);
      throw std::bad_alloc();
    }
    if (cnt > 0) {
      hipError_t e = hipMemcpyAsync(h, dout, sizeof(stq_pair) * cnt, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      (void)hipFree(dout);
      if (e != hipSuccess) {
        std::free(h);
        throw stqamd::HipError(e, "hipMemcpyAsync(D2H pairs)", __FILE__, __LINE__);
      }
    } else if (dout) {
      (void)hipFree(dout);
    }
    *h_out = h;
    *count = cnt;
  });
}

int stq_find_pairs_batched_device(void *ctx, const stq_aabb_f64 *d_boxes,
                                  const int64_t *h_scene_offsets, int32_t nscenes,
                                  stq_pair *d_out, int64_t capacity, int64_t *h_pair_offsets) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || nscenes < 0 || !h_scene_offsets || !h_pair_offsets || capacity < 0 ||
      (capacity > 0 && !d_out))
    return STQ_ERR_INVALID_ARGUMENT;
  for (int s = 0; s < nscenes; ++s) {
    const int64_t m = h_scene_offsets[s + 1] - h_scene_offsets[s];
    if (m < 0 || h_scene_offsets[s] < 0) return STQ_ERR_INVALID_ARGUMENT;
    if (m > INT32_MAX) return STQ_ERR_TOO_LARGE;
    if (m > 0 && !d_boxes) return STQ_ERR_INVALID_ARGUMENT;
  }
  return stqamd::guarded(c, [&] {
    const int64_t nb = nscenes > 0 ? h_scene_offsets[nscenes] - h_scene_offsets[0] : 0;
    stqamd::tuned_call(c, d_boxes, nb, h_scene_offsets, nscenes, true, [&] {
      c->run_batched(d_boxes, h_scene_offsets, nscenes, d_out, capacity, h_pair_offsets);
    });
  });
}

static unsigned *flag_word(Context *c) {
  return reinterpret_cast<unsigned *>(
      c->counters.as<unsigned long long>(stqamd::kCounters) + stqamd::kFlagCounter);
}

static void check_flag(Context *c, unsigned *err, const char *what) {
  unsigned h = 0;
  STQ_HIP(hipMemcpyAsync(&h, err, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  STQ_HIP(hipStreamSynchronize(c->stream));
  if (h) throw stqamd::StatusError(STQ_ERR_INVALID_ARGUMENT, what);
}

static bool mesh_args_valid(const double *v0, const double *v1, int64_t nv, const int32_t *e,
                            int64_t ne, const int32_t *f, int64_t nf) {
  if (nv < 0 || ne < 0 || nf < 0) return false;
  if (nv > 0 && (!v0 || !v1)) return false;
  if ((ne > 0 && !e) || (nf > 0 && !f)) return false;
  return true;
}

// Builds the boxes on the context's stream and checks the index flag.
// A call around the path (box build, orientation, query records): its stats are
// the record count and the kernel's time on the stream (ms_total), bracketed by
// HIP events; `launch` enqueues the kernel, then the error flag is read back.
extern "C++" template <class F>
static void timed_kernel(Context *c, int64_t n, unsigned *err, const char *what, F &&launch) {
  c->stats = stq_stats{};
  c->stats.n_boxes = n;
  STQ_HIP(hipMemsetAsync(err, 0, sizeof(unsigned), c->stream));
  STQ_HIP(hipEventRecord(c->ev[stqamd::E_START], c->stream));
  launch();
  STQ_HIP(hipEventRecord(c->ev[stqamd::E_END], c->stream));
  check_flag(c, err, what);
  c->stats.ms_total = stqamd::ev_ms(c->ev[stqamd::E_START], c->ev[stqamd::E_END]);
}

static void build_boxes(Context *c, const double *v0, const double *v1, int64_t nv,
                        const int32_t *e, int64_t ne, const int32_t *f, int64_t nf, double r,
                        bool float_build, stq_aabb_f64 *out) {
  timed_kernel(c, nv + ne + nf, flag_word(c), "edge or face index outside [0, nv)", [&] {
    stqamd::launch_build_boxes(v0, v1, nv, e, ne, f, nf, r, float_build, out, flag_word(c),
                               c->stream);
  });
}

int stq_build_swept_boxes_device(void *ctx, const double *d_v0, const double *d_v1, int64_t nv,
                                 const int32_t *d_edges, int64_t ne, const int32_t *d_faces,
                                 int64_t nf, double inflation_radius, int32_t float_build,
                                 stq_aabb_f64 *d_boxes) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !mesh_args_valid(d_v0, d_v1, nv, d_edges, ne, d_faces, nf) ||
      (nv + ne + nf > 0 && !d_boxes) || float_build < 0 || float_build > 1)
    return STQ_ERR_INVALID_ARGUMENT;
  if (nv + ne + nf > INT32_MAX || nv > INT32_MAX) return STQ_ERR_TOO_LARGE;
  return stqamd::guarded(c, [&] {
    build_boxes(c, d_v0, d_v1, nv, d_edges, ne, d_faces, nf, inflation_radius, float_build != 0,
                d_boxes);
  });
}

int stq_find_pairs_mesh_device(void *ctx, const double *d_v0, const double *d_v1, int64_t nv,
                               const int32_t *d_edges, int64_t ne, const int32_t *d_faces,
                               int64_t nf, double inflation_radius, stq_pair *d_out,
                               int64_t capacity, int64_t *count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !count || !mesh_args_valid(d_v0, d_v1, nv, d_edges, ne, d_faces, nf) || capacity < 0 ||
      (capacity > 0 && !d_out))
    return STQ_ERR_INVALID_ARGUMENT;
  const int64_t n = nv + ne + nf;
  if (n > INT32_MAX) return STQ_ERR_TOO_LARGE;
  return stqamd::guarded(c, [&] {
    stq_aabb_f64 *boxes = c->mesh_boxes.as<stq_aabb_f64>((size_t)std::max<int64_t>(n, 1));
    build_boxes(c, d_v0, d_v1, nv, d_edges, ne, d_faces, nf, inflation_radius, false, boxes);
    stqamd::tuned_call(c, boxes, n, nullptr, 1, false,
                       [&] { c->run(boxes, n, d_out, capacity, nullptr, count); });
  });
}

int stq_orient_pairs_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n, stq_pair *d_pairs,
                            int64_t count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || n < 0 || count < 0 || (n > 0 && !d_boxes) || (count > 0 && !d_pairs))
    return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    unsigned *err = flag_word(c);
    timed_kernel(c, count, err,
                 "pair of another kind than vertex-face or edge-edge, or id out of range",
                 [&] {
                   uint8_t *kinds = c->box_kinds.as<uint8_t>((size_t)std::max<int64_t>(n, 1));
                   stqamd::launch_box_kinds(d_boxes, n, kinds, c->stream);
                   stqamd::launch_orient_pairs(kinds, n, d_pairs, count, err, c->stream);
                 });
  });
}

int stq_ccd_queries_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n, const double *d_v0,
                           const double *d_v1, int64_t nv, const stq_pair *d_pairs, int64_t count,
                           double min_separation, stq_ccd_query *d_queries) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || n < 0 || nv < 0 || count < 0 || (n > 0 && !d_boxes) || (nv > 0 && (!d_v0 || !d_v1)) ||
      (count > 0 && (!d_pairs || !d_queries)))
    return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    unsigned *err = flag_word(c);
    timed_kernel(c, count, err,
                 "pair of another kind, id out of range, or vertex id outside [0, nv)", [&] {
                   uint8_t *kinds = c->box_kinds.as<uint8_t>((size_t)std::max<int64_t>(n, 1));
                   stqamd::launch_box_kinds(d_boxes, n, kinds, c->stream);
                   stqamd::launch_ccd_queries(d_boxes, kinds, n, d_v0, d_v1, nv, d_pairs, count,
                                              min_separation, d_queries, err, c->stream);
                 });
  });
}

int stq_sizeof_ccd_query(void) { return (int)sizeof(stq_ccd_query); }

int stq_autotune_save(void *ctx, const char *path) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !path) return STQ_ERR_INVALID_ARGUMENT;
  return c->tune_save(path);
}

int stq_autotune_load(void *ctx, const char *path) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !path) return STQ_ERR_INVALID_ARGUMENT;
  return c->tune_load(path);
}

int stq_autotune_clear(void *ctx) {
  Context *c = static_cast<Context *>(ctx);
  if (!c) return STQ_ERR_INVALID_ARGUMENT;
  c->tune_table.clear();
  return STQ_OK;
}

int stq_autotune_entries(void *ctx, int32_t *count) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !count) return STQ_ERR_INVALID_ARGUMENT;
  *count = (int32_t)c->tune_table.size();
  return STQ_OK;
}

int stq_device_alloc(void *ctx, size_t bytes, void **d_ptr) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !d_ptr) return STQ_ERR_INVALID_ARGUMENT;
  *d_ptr = nullptr;
  return stqamd::guarded(c, [&] {
    if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) {
      (void)hipGetLastError();
      *d_ptr = nullptr;
      throw stqamd::StatusError(STQ_ERR_OUT_OF_MEMORY, "device allocation failed");
    }
  });
}

int stq_copy_to_device(void *ctx, void *d_dst, const void *h_src, size_t bytes) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || (bytes && (!d_dst || !h_src))) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    if (!bytes) return;
    STQ_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_copy_to_host(void *ctx, void *h_dst, const void *d_src, size_t bytes) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || (bytes && (!d_src || !h_dst))) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    if (!bytes) return;
    STQ_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_free_device(void *ctx, void *d_ptr) {
  Context *c = static_cast<Context *>(ctx);
  if (!c) return STQ_ERR_INVALID_ARGUMENT;
  if (!d_ptr) return STQ_OK;
  return stqamd::guarded(c, [&] { STQ_HIP(hipFree(d_ptr)); });
}

void stq_free_host(void *h_ptr) { std::free(h_ptr); }

int stq_sizer_num_params(void) { return stqamd::kParams; }
int stq_sizer_num_features(void) { return stqamd::kFeat; }

int stq_sizer_get_params(void *ctx, float *h_params) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !h_params) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    c->sizer_ensure();
    STQ_HIP(hipMemcpyAsync(h_params, c->sz_params.p, sizeof(float) * stqamd::kParams,
                           hipMemcpyDeviceToHost, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_sizer_set_params(void *ctx, const float *h_params) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !h_params) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    c->sizer_ensure();
    STQ_HIP(hipMemcpyAsync(c->sz_params.p, h_params, sizeof(float) * stqamd::kParams,
                           hipMemcpyHostToDevice, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_sizer_reset(void *ctx, uint64_t seed) {
  Context *c = static_cast<Context *>(ctx);
  if (!c) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    c->sizer_ensure();
    std::vector<float> h(stqamd::kParams);
    stqamd::sizer_init_params(h.data(), seed);
    STQ_HIP(hipMemcpyAsync(c->sz_params.p, h.data(), sizeof(float) * stqamd::kParams,
                           hipMemcpyHostToDevice, c->stream));
    STQ_HIP(hipMemsetAsync(c->sz_m.p, 0, sizeof(float) * stqamd::kParams, c->stream));
    STQ_HIP(hipMemsetAsync(c->sz_v.p, 0, sizeof(float) * stqamd::kParams, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
    c->adam_t = 0;
    c->scenes_seen = 0;
    c->replay_n = 0;
    c->replay_pos = 0;
    c->train_calls = 0;
  });
}

// Host rows to device scratch (features and optional targets).
static void upload_rows(Context *c, const float *h_feat, const float *h_t, int64_t n, float **df,
                        float **dt) {
  *df = c->bfeat.as<float>((size_t)std::max<int64_t>(n, 1) * stqamd::kFeat);
  STQ_HIP(hipMemcpyAsync(*df, h_feat, sizeof(float) * stqamd::kFeat * n, hipMemcpyHostToDevice,
                         c->stream));
  if (dt) {
    *dt = c->btargets.as<float>((size_t)std::max<int64_t>(n, 1));
    STQ_HIP(hipMemcpyAsync(*dt, h_t, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  }
}

int stq_sizer_forward(void *ctx, const float *h_features, int64_t n, float *h_out) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || n < 0 || (n > 0 && (!h_features || !h_out))) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    if (n == 0) return;
    c->sizer_ensure();
    float *df = nullptr;
    upload_rows(c, h_features, nullptr, n, &df, nullptr);
    float *dp = c->pred.as<float>((size_t)n);
    stqamd::launch_mlp_forward(c->sz_params.as<float>(stqamd::kParams), df, n, dp, c->stream);
    STQ_HIP(hipMemcpyAsync(h_out, dp, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_sizer_gradient(void *ctx, const float *h_features, const float *h_targets, int64_t n,
                       float *h_grad, float *h_loss) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || n <= 0 || !h_features || !h_targets || !h_grad || !h_loss)
    return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    c->sizer_ensure();
    float *df = nullptr, *dt = nullptr;
    upload_rows(c, h_features, h_targets, n, &df, &dt);
    c->sizer_gradient(df, dt, n);
    STQ_HIP(hipMemcpyAsync(h_grad, c->sz_grad.p, sizeof(float) * stqamd::kParams,
                           hipMemcpyDeviceToHost, c->stream));
    STQ_HIP(hipMemcpyAsync(h_loss, c->sz_loss.p, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_sizer_train(void *ctx, const float *h_features, const float *h_targets, int64_t n,
                    int32_t steps, float lr) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || n <= 0 || steps < 0 || !h_features || !h_targets || !(lr > 0))
    return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    c->sizer_ensure();
    // rows go to the per-item buffers; training draws minibatches from them
    float *df = c->feat.as<float>((size_t)n * stqamd::kFeat);
    float *dt = c->targets.as<float>((size_t)n);
    STQ_HIP(hipMemcpyAsync(df, h_features, sizeof(float) * stqamd::kFeat * n, hipMemcpyHostToDevice,
                           c->stream));
    STQ_HIP(hipMemcpyAsync(dt, h_targets, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    c->sizer_train(df, dt, n, steps, lr);
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_sizer_last_items(void *ctx, float *h_features, uint32_t *h_counts, int64_t capacity,
                         int64_t *n) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !n || capacity < 0) return STQ_ERR_INVALID_ARGUMENT;
  return stqamd::guarded(c, [&] {
    const int64_t m = c->stats.n_tiles;
    *n = m;
    const int64_t k = std::min(m, capacity);
    if (k <= 0) return;
    if (h_features)
      STQ_HIP(hipMemcpyAsync(h_features, c->feat.p, sizeof(float) * stqamd::kFeat * k,
                             hipMemcpyDeviceToHost, c->stream));
    if (h_counts)
      STQ_HIP(hipMemcpyAsync(h_counts, c->counts.p, sizeof(uint32_t) * k, hipMemcpyDeviceToHost,
                             c->stream));
    STQ_HIP(hipStreamSynchronize(c->stream));
  });
}

int stq_get_stats(void *ctx, stq_stats *stats) {
  Context *c = static_cast<Context *>(ctx);
  if (!c || !stats) return STQ_ERR_INVALID_ARGUMENT;
  *stats = c->stats;
  return STQ_OK;
}

const char *stq_last_error(void *ctx) {
  Context *c = static_cast<Context *>(ctx);
  if (!c) return "null context";
  return c->last_error.c_str();
}

const char *stq_status_string(int status) {
  switch (status) {
    case STQ_OK: return "ok";
    case STQ_ERR_INVALID_ARGUMENT: return "invalid argument";
    case STQ_ERR_INVALID_BOX: return "invalid box (NaN or min > max)";
    case STQ_ERR_CAPACITY: return "output capacity too small";
    case STQ_ERR_OUT_OF_MEMORY: return "out of memory";
    case STQ_ERR_HIP: return "HIP runtime error";
    case STQ_ERR_TOO_LARGE: return "scene too large";
    case STQ_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

int stq_sizeof_aabb(void) { return (int)sizeof(stq_aabb_f64); }
int stq_sizeof_options(void) { return (int)sizeof(stq_options); }
int stq_sizeof_stats(void) { return (int)sizeof(stq_stats); }

}  // extern "C"
