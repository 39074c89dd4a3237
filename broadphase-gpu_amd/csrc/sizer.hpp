// sizer.hpp -- learned per-item pair-count predictor (sizer.hip).
//
// No counterpart in the reference, which sizes its output from free device
// memory and reruns on overflow (src/memory.cuh:163-188,
// src/simulation.cu:714-745), or counts then refills (src/simulation.cu:
// 451-489).  Here a small MLP predicts each sweep work item's pair count from
// 32 cheap features; the prediction sizes the item's slice of a scratch
// buffer for a single emitting pass, and items that overflow are redone
// exactly.
#pragma once

#include "stq_internal.hpp"
#include "sweep.hpp"

namespace stqamd {

constexpr int kFeat = 32;
constexpr int kHid = 64;
// W1[kHid][kFeat], b1[kHid], W2[kHid][kHid], b2[kHid], W3[kHid], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kFeat;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kParams = kOffB3 + 1;  // 6337

// Features of every work item (nitems x kFeat floats) and its exact number of
// sweep candidates (an upper bound on its pairs).  cols: the column statistics
// of the gather (ncols blocks of kColBlock positions).  One scene: qp and axis; batched scenes: sp/seg_off/nscenes
// give each item's scene, and the item count is sa.nitems_dev when set.
// ncols: the number of column blocks; col_work: col_work_doubles(ncols)
// doubles for their prefix sums.
void launch_features(bool f32, const SweepArgs &sa, const ColStats *cols, int64_t ncols,
                     double *col_work, const QParams &qp, int axis, const SceneParams *sp,
                     const int64_t *seg_off, int nscenes, float *feat, uint32_t *cand,
                     hipStream_t st);
size_t col_work_doubles(int64_t ncols);
// pred[i] = MLP(feat[i]) (the log1p of the predicted pair count).  Where an
// n_dev pointer is given, the count is *n_dev (n is then an upper bound).
void launch_mlp_forward(const float *params, const float *feat, int64_t n, float *pred,
                        hipStream_t st, const int64_t *n_dev = nullptr);
// caps[i] = min(cand[i], ceil(alpha * expm1(pred[i]) + beta)) for i < n
// (n = *n_dev when given; nothing is written beyond).
void launch_caps(const float *pred, const uint32_t *cand, int64_t n, float alpha, float beta,
                 uint32_t *caps, hipStream_t st, const int64_t *n_dev = nullptr);
// caps[i] = 0 where the item's slots [cap_off[i], cap_off[i] + caps[i]) end
// beyond `budget` (the scratch size): such items are redone exactly.
void launch_cap_budget(uint32_t *caps, const int64_t *cap_off, int64_t n, int64_t budget,
                       hipStream_t st, const int64_t *n_dev = nullptr);
// targets[i] = log1p(counts[i]).
void launch_targets(const uint32_t *counts, int64_t n, float *targets, hipStream_t st);
// Mean squared error of MLP(feat) against targets and its gradient (kParams
// floats), reduced in a fixed order; partial needs mlp_partial_floats(n).
void launch_mlp_gradient(const float *params, const float *feat, const float *targets, int64_t n,
                         float *partial, float *grad, float *loss, hipStream_t st);
size_t mlp_partial_floats(int64_t bs);
// One Adam step (t is the 1-based step number) on the minibatch of bs rows
// whose row k is source row ((start + k) * stride) mod n of feat/targets.
void launch_train_step(float *params, float *m, float *v, const float *feat, const float *targets,
                       int64_t n, int64_t start, int64_t stride, int64_t bs, float *partial, float lr,
                       int t, hipStream_t st);
// sum |expm1(pred) - counts| over items, added to *acc.
void launch_abs_error(const float *pred, const uint32_t *counts, int64_t n, double *acc,
                      hipStream_t st, const int64_t *n_dev = nullptr);
// Per item: if counts <= caps copy its pairs from scratch (at cap_off) to out
// (at off); otherwise append the item to ovf (count in *novf).  Nothing
// happens when *total exceeds capacity (total may be null).
void launch_compact(const uint32_t *counts, const uint32_t *caps, const int64_t *cap_off,
                    const int64_t *off, int64_t n, const int2 *scratch, int2 *out, int *ovf,
                    unsigned long long *novf, hipStream_t st, const int64_t *n_dev = nullptr,
                    const int64_t *total = nullptr, int64_t capacity = 0);

// Minibatch rows k = 0..bs-1 of item (start + k) * stride mod n.
void launch_gather_batch(const float *feat, const float *target, int64_t n, int64_t start,
                         int64_t stride, int64_t bs, float *bfeat, float *btarget, hipStream_t st);

// Deterministic initial parameters (He-uniform from a counter-based hash).
void sizer_init_params(float *host_params, uint64_t seed);

}  // namespace stqamd
