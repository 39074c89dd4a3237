// sweep.hip -- key build, sorted-record gather, chain ends, work planning and
// the sweep-and-tiniest-queue kernel for gfx950.
//
// Reference counterparts:
//   k_axis_stats  calc_mean / calc_variance (src/sweep.cu:93-123), disabled in
//                 the reference (src/simulation.cu:594-637); here FP64 with a
//                 fixed-order reduction so the axis choice is deterministic
//   k_keys_*      the key half of create_ds (src/sweep.cu:125-161)
//   k_gather_*    the MiniBox half of create_ds, after the sort
//   k_chain_end   the loop bound of twostage_queue: the chain of sorted row p
//                 is every q > p with min_q <= max_p (src/sweep.cu:245,269);
//                 sorted minima make it the range (p, hi_p]
//   k_plan        replaces the 64-slot shared Queue (src/queue.cuh:9-38) and
//                 the one-candidate-per-round loop (src/sweep.cu:252-278)
//   k_sweep       twostage_queue + does_collide / is_valid_pair / covertex +
//                 add_overlap (src/sweep.cu:226-279, src/collision.cu:38-74,
//                 src/aabb.cu:37-40)
//
// Sweep organisation.  A wave owns a group of up to 64 sorted rows, one per
// lane, and walks the candidate q positions of the group in order.  q is the
// same in every lane, so the q record arrives through scalar loads into SGPRs
// and each comparison is one VALU op over 64 rows.  Rows are regrouped per
// window of W consecutive sorted rows by their chain end hi_p, so the 64
// rows of a group end their chains close together; each group's q range is
// cut into chunks of at most `chunk` positions, and each (group, chunk) is a
// work item taken by the persistent grid.  Pair appends are wave aggregated:
// one ballot, one popcount and per-lane mbcnt offsets; the count pass stores
// one number per item, the fill pass writes at the item's prefix-sum offset,
// so the output order is fixed by the input.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "plan.hpp"
#include "stq_internal.hpp"
#include "sweep.hpp"

namespace stqamd {

namespace {

constexpr int kThreads = 256;

__device__ inline unsigned mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ inline bool bad(double a, double b) { return !(a <= b); }  // NaN or inverted

// ---- axis statistics ----------------------------------------------------------
// Per axis: sum(min+max), sum(min^2+max^2), sum(max-min), sum(((min+max)/2)^2),
// then min(min) and max(max) for the quantisation range.
__global__ __launch_bounds__(kThreads) void k_axis_stats(const stq_aabb_f64 *__restrict__ boxes,
                                                         int64_t n, double *__restrict__ partial,
                                                         uint32_t *__restrict__ keys3,
                                                         unsigned *__restrict__ err) {
  double acc[12], lo[3], up[3];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = INFINITY;
    up[a] = -INFINITY;
  }
  if (!keys3) {
    axis_stats_range(boxes, (int64_t)blockIdx.x * kThreads + threadIdx.x, n,
                     (int64_t)gridDim.x * kThreads, acc, lo, up);
  } else {
    bool bad = false;
    axis_stats_keys_range(boxes, (int64_t)blockIdx.x * kThreads + threadIdx.x, n,
                          (int64_t)gridDim.x * kThreads, acc, lo, up, keys3, n, bad);
    if (bad) atomicOr(err, kErrInverted);
  }
  block_axis_stats<kThreads>(acc, lo, up, partial + (int64_t)blockIdx.x * kAxisStats);
}

// One block: the totals of the per-block statistics (combined in a fixed
// order, so the result does not depend on timing), the quantisation range of
// each axis and the sweep axis of the policy:
//   0..2 that axis; 3 (variance) the largest sum over end points of
//   (x - mu)^2, mu the mean centre (src/sweep.cu:93-123); 4 (chain) the
//   smallest mean extent over the spread of the centres.
// kLevelMax levels cover each axis's range; a degenerate or unbounded range
// puts every box on the full level range (scale 0).
__global__ __launch_bounds__(kThreads) void k_axis_final(const double *__restrict__ partial,
                                                         int nb, int64_t n, int policy, int kbits,
                                                         SceneParams *__restrict__ out) {
  __shared__ double tot[kAxisStats];
  // thread t combines blocks t, t + kThreads, ... in order, then the threads
  // are combined in a fixed order (the same on every call)
  double acc[12], lo[3], up[3];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = INFINITY;
    up[a] = -INFINITY;
  }
  for (int b = threadIdx.x; b < nb; b += kThreads) {
    const double *x = partial + (int64_t)b * kAxisStats;
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] += x[k];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], x[12 + a]);
      up[a] = fmax(up[a], x[15 + a]);
    }
  }
  block_axis_stats<kThreads>(acc, lo, up, tot);
  __syncthreads();
  if (threadIdx.x != 0) return;
  SceneParams sp;
  for (int a = 0; a < 3; ++a) {
    const double lo = tot[12 + a], up = tot[15 + a], range = up - lo;
    sp.qp.origin[a] = isfinite(lo) ? lo : 0.0;
    const double sc = (range > 0 && isfinite(range)) ? kLevelMax / range : 0.0;
    sp.qp.scale[a] = isfinite(sc) ? sc : 0.0;
  }
  int best = policy <= 2 ? policy : 0;
  if (policy > 2) {
    const double dn = (double)n;
    double bestv = 0;
    for (int a = 0; a < 3; ++a) {
      const double s1 = tot[4 * a + 0], s2 = tot[4 * a + 1], ext = tot[4 * a + 2],
                   cc = tot[4 * a + 3];
      const double mu = s1 / (2.0 * dn);
      double v;
      if (policy == STQ_AXIS_VARIANCE) {
        v = s2 - 2.0 * mu * s1 + 2.0 * dn * mu * mu;
        if (a == 0 || v > bestv) {
          best = a;
          bestv = v;
        }
      } else {
        const double sd = sqrt(fmax(cc / dn - mu * mu, 0.0));
        v = sd > 0 ? (ext / dn) / sd : INFINITY;
        if (a == 0 || v < bestv) {
          best = a;
          bestv = v;
        }
      }
    }
  }
  sp.axis = best;
  set_sweep_keys(sp, tot[12 + best], tot[15 + best], kbits);
  *out = sp;
}

// ---- the bulk of each scene's sweep-key range ------------------------------------
// One block per scene: a strided sample of at most kRangeSample boxes, its
// minima and maxima sorted together in LDS (one bitonic network, padded with
// infinities), then the range [qlo - w/8, qhi + w/8] with w = qhi - qlo,
// kept inside the scene's finite range.
constexpr int kRangeSample = 512;
__global__ __launch_bounds__(kThreads) void k_sweep_range(const stq_aabb_f64 *__restrict__ boxes,
                                                          const int64_t *__restrict__ seg_off,
                                                          int64_t n,
                                                          SceneParams *__restrict__ sp) {
  __shared__ double smn[kRangeSample], smx[kRangeSample];
  const int sc = blockIdx.x;
  const int64_t b0 = seg_off ? seg_off[sc] : 0, m = seg_off ? seg_off[sc + 1] - b0 : n;
  if (m < 2) return;
  const int axis = sp[sc].axis;
  const int S = (int)(m < kRangeSample ? m : kRangeSample);
  for (int i = threadIdx.x; i < kRangeSample; i += kThreads) {
    double a = INFINITY, c = -INFINITY;  // padding: after the minima, before the maxima
    if (i < S) {
      const stq_aabb_f64 &b = boxes[b0 + (int64_t)i * m / S];
      a = b.min[axis];
      c = b.max[axis];
    }
    smn[i] = a;
    smx[i] = c;
  }
  __syncthreads();
  for (int size = 2; size <= kRangeSample; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int j = threadIdx.x; j < kRangeSample; j += kThreads) {
        const int pj = j ^ stride;
        if (pj > j) {
          const bool up = (j & size) == 0;
          const double a = smn[j], c = smn[pj];
          if ((a > c) == up) {
            smn[j] = c;
            smn[pj] = a;
          }
          const double x = smx[j], y = smx[pj];
          if ((x > y) == up) {
            smx[j] = y;
            smx[pj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  const int k = S / 200;  // 0.5% of the sample on each side
  if (threadIdx.x == 0) {
    const double qlo = smn[k], qhi = smx[kRangeSample - 1 - k];
    const double w = qhi - qlo;
    const SceneParams p = sp[sc];  // the finite range from the statistics
    const double flo = p.korigin, fup = p.kscale > 0 ? p.korigin + (double)p.kmax / p.kscale : flo;
    const double lo = fmax(qlo - 0.125 * w, flo), up = fmin(qhi + 0.125 * w, fup);
    const double sc2 = (double)sp[sc].kmax / (up - lo);
    if (isfinite(lo) && isfinite(up) && up > lo && isfinite(sc2)) {
      sp[sc].korigin = lo;
      sp[sc].kscale = sc2;
    }
  }
}

// ---- keys ---------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_keys_f64(const stq_aabb_f64 *__restrict__ boxes,
                                                       int64_t n, int axis,
                                                       const SceneParams *__restrict__ sp,
                                                       uint64_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals,
                                                       unsigned *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  if (sp) axis = sp->axis;
  const stq_aabb_f64 &b = boxes[i];
  const double x0 = b.min[0], x1 = b.min[1], x2 = b.min[2];
  const double y0 = b.max[0], y1 = b.max[1], y2 = b.max[2];
  if (bad(x0, y0) || bad(x1, y1) || bad(x2, y2)) atomicOr(err, kErrInverted);
  const double m = axis == 0 ? x0 : (axis == 1 ? x1 : x2);
  keys[i] = key_of(m);
  vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kThreads) void k_keys_f32(const stq_aabb_f64 *__restrict__ boxes,
                                                       int64_t n, int axis,
                                                       const SceneParams *__restrict__ sp,
                                                       uint32_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals,
                                                       unsigned *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  if (sp) axis = sp->axis;
  const stq_aabb_f64 &b = boxes[i];
  const double x0 = b.min[0], x1 = b.min[1], x2 = b.min[2];
  const double y0 = b.max[0], y1 = b.max[1], y2 = b.max[2];
  if (bad(x0, y0) || bad(x1, y1) || bad(x2, y2)) atomicOr(err, kErrInverted);
  const double m = axis == 0 ? x0 : (axis == 1 ? x1 : x2);
  keys[i] = key_of(round_down_f32(m));
  vals[i] = (uint32_t)i;
}

// ---- gather sorted records --------------------------------------------------
__device__ inline void other_axes(int axis, int &a1, int &a2) {
  a1 = axis == 0 ? 1 : 0;
  a2 = axis == 2 ? 1 : 2;
}

// Sorted records of position s; with cols, the column statistics of every
// kColBlock positions as well.
__global__ __launch_bounds__(kThreads) void k_gather_f64(const stq_aabb_f64 *__restrict__ boxes,
                                                         const uint32_t *__restrict__ order,
                                                         int64_t n, int axis,
                                                         const SceneParams *__restrict__ sp,
                                                         double *__restrict__ smin,
                                                         double *__restrict__ smax,
                                                         Rec64 *__restrict__ rec,
                                                         QRec *__restrict__ qrec,
                                                         ColStats *__restrict__ cols,
                                                         uint2 *__restrict__ rowf) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const bool valid = s < n;
  QParams qp{};
  if (sp) {
    axis = sp->axis;
    qp = sp->qp;
  }
  int a1, a2;
  other_axes(axis, a1, a2);
  Rec64 r{};
  if (valid) {
    const stq_aabb_f64 &b = boxes[order[s]];
    smin[s] = b.min[axis];
    smax[s] = b.max[axis];
    r.lo1 = b.min[a1];
    r.hi1 = b.max[a1];
    r.lo2 = b.min[a2];
    r.hi2 = b.max[a2];
    r.v0 = b.vertex_ids[0];
    r.v1 = b.vertex_ids[1];
    r.v2 = b.vertex_ids[2];
    r.type = type_bit(r.v1, r.v2);
    rec[s] = r;
    // levels for the axis-1 grouping of the grouped FP64 form
    if (qrec)
      qrec[s] = make_qrec(r.lo1, r.hi1, r.lo2, r.hi2, r.type, qp.origin[a1], qp.scale[a1],
                          qp.origin[a2], qp.scale[a2]);
    if (rowf)
      rowf[s] = row_feat(r.lo1, r.hi1, r.lo2, r.hi2, r.type, qp.origin[a1],
                         qp.scale[a1] / kLevelMax, qp.origin[a2], qp.scale[a2] / kLevelMax);
  }
  if (cols)
    col_stats_put(valid, s, r.lo1, r.hi1, r.lo2, r.hi2, r.type, qp.origin[a1],
                  qp.scale[a1] / kLevelMax, qp.origin[a2], qp.scale[a2] / kLevelMax, cols);
}

__global__ __launch_bounds__(kThreads) void k_gather_f32(const stq_aabb_f64 *__restrict__ boxes,
                                                         const uint32_t *__restrict__ order,
                                                         int64_t n, const SceneParams *__restrict__ sp,
                                                         uint32_t *__restrict__ smin,
                                                         uint32_t *__restrict__ smax,
                                                         Exact64 *__restrict__ rec,
                                                         QRec *__restrict__ qrec,
                                                         ColStats *__restrict__ cols,
                                                         uint2 *__restrict__ rowf) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const bool valid = s < n;
  const SceneParams p = *sp;
  const int axis = p.axis;
  const QParams qp = p.qp;
  int a1, a2;
  other_axes(axis, a1, a2);
  Exact64 r{};
  if (valid) {
    const stq_aabb_f64 &b = boxes[order[s]];
    smin[s] = skey_down(round_down_f32(b.min[axis]), p);  // = the sort key
    smax[s] = skey_up(round_up_f32(b.max[axis]), p);
    r.lo0 = b.min[axis];
    r.hi0 = b.max[axis];
    r.lo1 = b.min[a1];
    r.hi1 = b.max[a1];
    r.lo2 = b.min[a2];
    r.hi2 = b.max[a2];
    r.v0 = b.vertex_ids[0];
    r.v1 = b.vertex_ids[1];
    r.v2 = b.vertex_ids[2];
    r.type = type_bit(r.v1, r.v2);
    rec[s] = r;
    qrec[s] = make_qrec(r.lo1, r.hi1, r.lo2, r.hi2, r.type, qp.origin[a1], qp.scale[a1],
                        qp.origin[a2], qp.scale[a2]);
    if (rowf)
      rowf[s] = row_feat(r.lo1, r.hi1, r.lo2, r.hi2, r.type, qp.origin[a1],
                         qp.scale[a1] / kLevelMax, qp.origin[a2], qp.scale[a2] / kLevelMax);
  }
  if (cols)
    col_stats_put(valid, s, r.lo1, r.hi1, r.lo2, r.hi2, r.type, qp.origin[a1],
                  qp.scale[a1] / kLevelMax, qp.origin[a2], qp.scale[a2] / kLevelMax, cols);
}

// ---- chain ends -----------------------------------------------------------------
// hi[s] = last q with smin[q] <= smax[s] (q >= s since smin[s] <= smax[s]),
// by binary searches in tiles of rows whose minima are staged in LDS
// (chain_end_tile); a persistent grid of two blocks per CU adds sum(hi - s)
// to *cand with one atomic per block.  Rows outside [r0, r1) get an empty chain
// (hi = s): they emit nothing, as the reference's rows outside
// [tidstart, tidstart + MAX_OVERLAP_CUTOFF) (src/sweep.cu:234-240).
template <class T>
__global__ __launch_bounds__(kCeThreads, 8) void k_chain_end(const T *__restrict__ smin,
                                                          const T *__restrict__ smax, int64_t n,
                                                          int64_t r0, int64_t r1,
                                                          int *__restrict__ hi,
                                                          unsigned long long *__restrict__ cand) {
  __shared__ __align__(16) T win[kCeWin<T>];  // 16-byte window stores
  unsigned long long len = 0;
  const int64_t ntiles = (n + kCeRows<T> - 1) / kCeRows<T>;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x)
    len += chain_end_tile(smin, smax, n, t * kCeRows<T>, r0, r1, win,
                          [n](int64_t) { return n; }, hi);
  block_atomic_add<kCeThreads>(len, cand);
}

// ---- work planning ----------------------------------------------------------------
// Blocks take windows of W sorted rows (W a power of two) in grid-stride
// order.  The rows of a window are ordered with a bitonic sort in LDS, by
// (chain end, position), or with qrec (Q) by (group_key, position), and cut
// into groups of 64; with qrec and tails on, each group's rows are then put
// in chain-end order by a wave sort (the tail cut needs that order; without
// tails a group's rows stay in group_key order).  Group g covers rowlist[64g .. 64g+nrows); plan_group
// splits it into a body over [min p + 1, cut] and, for rows with much longer
// chains, a tail.  nchunks[g] = number of items of the group.
template <bool Q>
__global__ __launch_bounds__(kPlanThreads) void k_plan(const int *__restrict__ hi,
                                                   const QRec *__restrict__ qrec, int64_t n,
                                                   int W, int chunk, float tail_cost, bool cap,
                                                   int2 *__restrict__ rowlist,
                                                   int4 *__restrict__ groups,
                                                   int4 *__restrict__ tgroups,
                                                   int2 *__restrict__ gchunk,
                                                   uint32_t *__restrict__ nchunks,
                                                   unsigned long long *__restrict__ lane_steps,
                                                   const uint2 *__restrict__ rowf,
                                                   uint2 *__restrict__ rowfeat) {
  typedef typename std::conditional<Q, uint32_t, uint64_t>::type K;
  __shared__ K key[Q ? kMaxWindow32 : kMaxWindow64];
  __shared__ unsigned rs_cnt[Q ? (kPlanThreads / 64) * 256 : 1], rs_sum[Q ? 256 : 1];
  const int64_t nwin = (n + W - 1) / W;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long steps = 0;
  for (int64_t b = blockIdx.x; b < nwin; b += gridDim.x) {  // grid-stride over windows
    const int64_t wbase = b * W;
    const int cnt = (int)std::min<int64_t>(W, n - wbase);
    plan_window_keys<Q, K>(hi, qrec, wbase, cnt, W, key);
    if constexpr (Q) {
      if (W >= kPlanThreads)
        window_radix_sort<kPlanThreads>(key, W, rs_cnt, rs_sum);
      else
        window_sort_any<kPlanThreads, kMaxWindow32>(key, W);
    } else {
      window_sort_any<kPlanThreads, kMaxWindow64>(key, W);
    }
    const int ng = (cnt + 63) / 64;
    const bool tails = Q && tail_cost > 0.f;
    // with qrec, the rows' window offsets go to the radix counts' LDS (free
    // again) and their chain ends over their keys
    uint16_t *off16 = reinterpret_cast<uint16_t *>(rs_cnt);
    int *hend = reinterpret_cast<int *>(key);
    if constexpr (Q)
      plan_rows_q(hi, rowf, tails, reinterpret_cast<const uint32_t *>(key), off16, hend, wbase,
                  cnt, ng, w, lane, rowlist, rowfeat);
    // one wave per group
    for (int k = w; k < ng; k += kPlanThreads / 64) {
      const int j = 64 * k + lane;
      const bool act = j < cnt;
      int p = INT_MAX, h = INT_MIN;
      if (act) {
        if constexpr (Q) {
          p = (int)(wbase + off16[j]);
          h = hend[j];
        } else {  // the key is (chain end, window offset)
          p = (int)(wbase + (uint32_t)key[j]);
          h = (int)(key[j] >> 32);
        }
      }
      if (tails) {  // chain-end order, which only the tail cut needs
        const uint64_t r = wave_sort_u64(act ? ((uint64_t)(uint32_t)h << 32) | (uint32_t)p : ~0ull,
                                         lane);
        if (act) {
          p = (int)(uint32_t)r;
          h = (int)(r >> 32);
        }
      }
      if (act && (tails || !Q)) {
        rowlist[wbase + j] = make_int2(p, h);
        if (rowf) rowfeat[wbase + j] = rowf[p];
      }
      const int nrows = min(64, cnt - 64 * k);
      const GroupPlan gp = plan_group(lane, nrows, p, h, (int)(wbase + 64 * k), chunk, cap,
                                      tail_cost);
      const int64_t g = (wbase >> 6) + k;
      if (lane == 0) {
        groups[g] = gp.body;
        tgroups[g] = gp.tail;
        gchunk[g] = gp.chunk;
        nchunks[g] = gp.nitems;
        steps += gp.steps;
      }
    }
    __syncthreads();  // the next window reuses key[]
  }
  block_atomic_add<kPlanThreads>(steps, lane_steps);
}

// Items of every group: the body's chunks, then the tail's (items[ioff[g]..]).
__global__ __launch_bounds__(kThreads) void k_items(const int4 *__restrict__ groups,
                                                    const int4 *__restrict__ tgroups,
                                                    const int2 *__restrict__ gchunk,
                                                    const int64_t *__restrict__ ioff, int64_t ngroups,
                                                    Item *__restrict__ items) {
  const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (g >= ngroups) return;
  write_group_items(groups[g], tgroups[g], gchunk[g], ioff[g], ioff[g + 1] - ioff[g], items);
}

// ---- work split by predicted cost ----------------------------------------------------
// Cost of an item in units of about 40 ns of one wave, from a least-squares
// fit of measured per-item times of the reduced-precision sweep (about 4.6 us
// per item, 0.41 us per 64 positions, 0.07 us per row and 0.04 us per pair):
// the grouped sweeps scan 64 positions per step; the plain FP64 sweep one
// position per step (about a third of a 64-position step); a tail item tests
// each of its rows against each 64 positions.
__global__ __launch_bounds__(kThreads) void k_item_cost(const Item *__restrict__ items, int64_t n,
                                                        const int64_t *__restrict__ n_dev,
                                                        bool grouped,
                                                        const float *__restrict__ pred,
                                                        const uint32_t *__restrict__ cand,
                                                        const uint32_t *__restrict__ counts,
                                                        uint32_t *__restrict__ cost) {
  const int64_t m = n_dev ? std::min(*n_dev, n) : n;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kThreads) {
    const Item it = items[i];
    const int rows = it.nrows & kRowsMask;
    const int64_t span = std::max<int64_t>(0, (int64_t)it.q1 - it.q0 + 1);
    const int64_t blocks = (span + 63) >> 6;
    // the network predicts log(1 + pairs); an item has at most cand[i]
    int64_t pairs = counts ? (int64_t)counts[i] : 0;
    if (pred) {
      const float p = expm1f(fminf(fmaxf(pred[i], 0.f), 30.f));
      pairs = (int64_t)fminf(p, cand ? (float)cand[i] : 1e9f);
    }
    int64_t c;
    if (it.nrows & kTailFlag)
      c = 64 + blocks * rows + pairs;
    else if (grouped)
      c = 64 + 10 * blocks + 2 * rows + pairs;
    else
      c = 64 + 3 * span + 2 * rows + pairs;
    cost[i] = (uint32_t)std::min<int64_t>(c, int64_t(1) << 30);
  }
}

// slices[x] = first item whose exclusive cost prefix reaches x / kXcds of the
// total (x = 0 .. kXcds); slices[kXcds] = the item count.
__global__ void k_xcd_slices(const int64_t *__restrict__ prefix, int64_t n,
                             const int64_t *__restrict__ n_dev, int64_t *__restrict__ slices) {
  const int x = threadIdx.x;
  if (x > kXcds) return;
  const int64_t m = n_dev ? std::min(*n_dev, n) : n;
  const int64_t total = prefix[m];
  const int64_t target = total / kXcds * x + total % kXcds * x / kXcds;
  int64_t lo = 0, hi = m;  // lower bound over prefix[0..m]
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (prefix[mid] < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  slices[x] = x == kXcds ? m : lo;
}

// ---- the sweep ----------------------------------------------------------------------
// Every test is a v_cmp whose lane mask is combined with s_and: the ballot of
// each comparison is taken separately, so no short-circuit branches appear.
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Work split of the persistent grid.  Blocks are dispatched to the XCDs round
// robin (block b on XCD b % kXcds); with xcd_split each XCD takes one
// contiguous slice of the item list, so its L2 holds only that slice's q
// records.
// With work counters (one per XCD slice) the waves claim items one at a time,
// so a wave that drew short items takes more of them; otherwise wave w takes
// items w, w + nwaves, ...
struct WaveSlice {
  int64_t kbeg, kend;
  int wave, nwaves, xcd;
};
__device__ __forceinline__ WaveSlice wave_slice(int64_t nitems, int xcd_split,
                                               const int64_t *slices) {
  WaveSlice w;
  w.xcd = 0;
  if (xcd_split) {
    const int xcd = (int)(blockIdx.x % kXcds);
    w.xcd = xcd;
    const int per = (int)(gridDim.x / kXcds);
    w.wave = uniform((int)((blockIdx.x / kXcds) * (kThreads / 64) + (threadIdx.x >> 6)));
    w.nwaves = per * (kThreads / 64);
    if (slices) {
      w.kbeg = slices[xcd];
      w.kend = slices[xcd + 1];
    } else {
      w.kbeg = nitems * xcd / kXcds;
      w.kend = nitems * (xcd + 1) / kXcds;
    }
  } else {
    w.wave = uniform((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
    w.nwaves = (int)gridDim.x * (kThreads / 64);
    w.kbeg = 0;
    w.kend = nitems;
  }
  return w;
}

// Claims the next item of the slice: lane 0's atomic result, read out only
// when the item is needed (claim_issue at the start of the previous item,
// claim_take at its end), so the atomic's latency overlaps the sweep.  Claim
// v is item kbeg + nwaves + v: items kbeg .. kbeg + nwaves - 1 are the waves'
// first ones, taken without a claim, so a launch starts without a burst of
// atomics and a wave with an empty slice issues none.
__device__ __forceinline__ unsigned claim_issue(unsigned *ctr) {
  unsigned v = 0;
  if ((threadIdx.x & 63) == 0) v = atomicAdd(ctr, 1u);
  return v;
}
__device__ __forceinline__ int64_t claim_take(const WaveSlice &ws, unsigned v) {
  return ws.kbeg + ws.nwaves + (int64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)v);
}

// Bounds of the unchecked middle part: every active lane has p < q <= hi_p
// for q in (pmax, hmin].
__device__ __forceinline__ void row_span(bool act, int p, int myhi, int &pmax, int &hmin) {
  pmax = act ? p : INT_MIN;
  hmin = act ? myhi : INT_MAX;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    pmax = max(pmax, __shfl_xor(pmax, o, kWave));
    hmin = min(hmin, __shfl_xor(hmin, o, kWave));
  }
  pmax = uniform(pmax);
  hmin = uniform(hmin);
}

// Emission of the pairs of the rows in hm (wave aggregated), as canonical
// (min, max) pairs of input ids: op is the row's input id, oq that of the
// partner (both unused when counting).
template <int MODE>
struct Emit {
  unsigned cnt = 0;
  int64_t base0 = 0, base = 0;
  unsigned cap = 0;
  __device__ __forceinline__ void put(bool h, uint64_t hm, uint32_t op, uint32_t oq, int2 *out) {
    const int2 pr = make_int2((int)min(op, oq), (int)max(op, oq));
    if (MODE == kFill) {
      if (h) out[base + mbcnt64(hm)] = pr;
      base += __popcll(hm);
    } else if (MODE == kLearned) {
      const unsigned slot = cnt + mbcnt64(hm);
      if (h && slot < cap) out[base0 + slot] = pr;
      cnt += (unsigned)__popcll(hm);
    } else {
      cnt += (unsigned)__popcll(hm);
    }
  }
};

// ---- plain FP64 sweep ----
// q records (Rec64, 48 B) arrive by scalar loads B at a time.  Scalar loads
// return out of order, so a wait for one batch waits for every scalar load in
// flight; the next batch's address is made to depend on the current batch
// (times the runtime zero) so its loads are issued after that wait and stay in
// flight while the current batch is tested.  Two buffers alternate.
template <class R>
__device__ __forceinline__ uint64_t geo_mask(const R &m, const R &q) {
  return ballot(m.hi1 >= q.lo1) & ballot(m.lo1 <= q.hi1) & ballot(m.hi2 >= q.lo2) &
         ballot(m.lo2 <= q.hi2);
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// ---- tail items ----
// A tail item holds up to 64 rows (the longest chains of a group, in order of
// chain end) and a position range [q0, q1].  Positions go across the lanes,
// 64 at a time; each row in turn is tested against them, its record broadcast
// from the lane that holds it.  A row whose chain ends before the current 64
// positions is skipped from then on.  Emission order: positions block, row,
// lane.
template <int MODE>
__device__ __forceinline__ void tail64(const SweepArgs &a, const Rec64 *__restrict__ rec,
                                       int2 *__restrict__ out, const Item &item, int lane,
                                       Emit<MODE> &em) {
  const int R = item.nrows & kRowsMask;
  int rp = INT_MAX, rh = INT_MIN;
  Rec64 rr;
  rr.lo1 = rr.lo2 = INFINITY;
  rr.hi1 = rr.hi2 = -INFINITY;
  rr.v0 = rr.v1 = rr.v2 = rr.type = 0;
  uint32_t rop = 0;
  if (lane < R) {
    const int2 r = a.rowlist[item.rowoff + lane];
    rp = r.x;
    rh = r.y;
    rr = rec[rp];
    if (MODE != kCount) rop = a.orig[rp];
  }
  int r0 = 0;
  for (int qb = item.q0; qb <= item.q1; qb += kWave) {
    while (r0 < R && readlane_i(rh, r0) < qb) ++r0;
    const int q = qb + lane;
    const bool inq = q <= item.q1;
    Rec64 mq;
    mq.lo1 = mq.lo2 = INFINITY;
    mq.hi1 = mq.hi2 = -INFINITY;
    mq.v0 = mq.v1 = mq.v2 = mq.type = 0;
    uint32_t oq = 0;
    if (inq) {
      mq = rec[q];
      if (MODE != kCount) oq = a.orig[q];
    }
    for (int r = r0; r < R; ++r) {
      const int p = readlane_i(rp, r), hp = readlane_i(rh, r);
      const double lo1 = readlane_f64(rr.lo1, r), hi1 = readlane_f64(rr.hi1, r);
      const double lo2 = readlane_f64(rr.lo2, r), hi2 = readlane_f64(rr.hi2, r);
      const int v0 = readlane_i(rr.v0, r), v1 = readlane_i(rr.v1, r), v2 = readlane_i(rr.v2, r);
      const int acc = accept_bits(readlane_i(rr.type, r));
      const bool h = inq & (q > p) & (q <= hp) & (mq.lo1 <= hi1) & (mq.hi1 >= lo1) &
                     (mq.lo2 <= hi2) & (mq.hi2 >= lo2) & ((acc & mq.type) != 0) &
                     (v0 != mq.v0) & (v0 != mq.v1) & (v0 != mq.v2) & (v1 != mq.v0) &
                     (v1 != mq.v1) & (v1 != mq.v2) & (v2 != mq.v0) & (v2 != mq.v1) &
                     (v2 != mq.v2);
      const uint64_t hm = ballot(h);
      if (hm) em.put(h, hm, (uint32_t)readlane_i((int)rop, r), oq, out);
    }
  }
}

// The record and output pointers are separate restrict-qualified arguments:
// that is what lets the compiler keep the uniform record loads on the scalar
// path next to the pair stores.
template <int MODE, int B>
__global__ __launch_bounds__(kThreads) void k_sweep64(SweepArgs a, const Rec64 *__restrict__ rec,
                                                      int2 *__restrict__ out) {
  const int zero = a.zero;
  const int lane = threadIdx.x & 63;
  if (MODE == kFill && a.total_dev && *a.total_dev > a.capacity) return;
  const WaveSlice ws = wave_slice(a.nitems_dev ? *a.nitems_dev : a.nitems, a.xcd_split, a.slices);
  // the first item is the wave's own; the claims number the rest from nwaves
  unsigned *wc = a.work ? a.work + ws.xcd * kClaimStride : nullptr;
  int64_t k = ws.kbeg + ws.wave;
  while (k < ws.kend) {
    const unsigned pend = wc ? claim_issue(wc) : 0u;
    const int64_t it = a.sel ? (int64_t)a.sel[k] : k;
    const Item item = a.items[it];
    Emit<MODE> em;
    if (MODE != kCount) em.base0 = em.base = a.offsets[it];
    if (MODE == kLearned) em.cap = a.caps[it];
    if (item.nrows & kTailFlag) {
      tail64<MODE>(a, rec, out, item, lane, em);
    } else {
      const bool act = lane < item.nrows;
      const int2 row = act ? a.rowlist[item.rowoff + lane] : make_int2(INT_MAX, INT_MIN);
      const int p = row.x, myhi = row.y;
      Rec64 mine;
      mine.lo1 = mine.lo2 = INFINITY;  // empty box for idle lanes
      mine.hi1 = mine.hi2 = -INFINITY;
      mine.v0 = mine.v1 = mine.v2 = mine.type = 0;
      uint32_t op = 0;
      if (act) {
        mine = rec[p];
        if (MODE != kCount) op = a.orig[p];
      }
      const int acc = act ? accept_bits(mine.type) : 0;

      // rows in m passed the geometric test against sorted q (record rq); idle
      // lanes can pass it only against unbounded boxes and never the type test
      auto hit = [&](int q, const Rec64 &rq, uint64_t m) __attribute__((always_inline)) {
        const bool h = (((m >> lane) & 1ull) != 0) & ((acc & rq.type) != 0) &
                       (mine.v0 != rq.v0) & (mine.v0 != rq.v1) & (mine.v0 != rq.v2) &
                       (mine.v1 != rq.v0) & (mine.v1 != rq.v1) & (mine.v1 != rq.v2) &
                       (mine.v2 != rq.v0) & (mine.v2 != rq.v1) & (mine.v2 != rq.v2);
        const uint64_t hm = ballot(h);
        if (hm) em.put(h, hm, op, MODE != kCount ? a.orig[q] : 0u, out);
      };
      auto batch = [&](const Rec64 *b, int q, bool check) __attribute__((always_inline)) {
        uint64_t m[B];
        uint64_t any = 0;
  #pragma unroll
        for (int j = 0; j < B; ++j) {
          m[j] = geo_mask(mine, b[j]);
          if (check) m[j] &= ballot(q + j > p) & ballot(q + j <= myhi);
          any |= m[j];
        }
        if (any) {
  #pragma unroll
          for (int j = 0; j < B; ++j)
            if (m[j]) hit(q + j, b[j], m[j]);
        }
      };
      auto walk = [&](int qa, int qb, bool check) __attribute__((always_inline)) {
        int q = qa;
        int nb = (qb - qa + 1) / B;
        if (nb > 0) {
          Rec64 b0[B], b1[B];
  #pragma unroll
          for (int j = 0; j < B; ++j) b0[j] = rec[q + j];
          while (true) {
            q += zero * b0[0].v0;  // waits for b0 on every path
            if (nb > 1) {
  #pragma unroll
              for (int j = 0; j < B; ++j) b1[j] = rec[q + B + j];
            }
            batch(b0, q, check);
            q += B;
            if (--nb == 0) break;
            q += zero * b1[0].v0;
            if (nb > 1) {
  #pragma unroll
              for (int j = 0; j < B; ++j) b0[j] = rec[q + B + j];
            }
            batch(b1, q, check);
            q += B;
            if (--nb == 0) break;
          }
        }
        for (; q <= qb; ++q) {
          const Rec64 r = rec[q];
          uint64_t m = geo_mask(mine, r);
          if (check) m &= ballot(q > p) & ballot(q <= myhi);
          if (m) hit(q, r, m);
        }
      };
      int pmax, hmin;
      row_span(act, p, myhi, pmax, hmin);
      const int q0 = item.q0, q1 = item.q1;
      const int b0 = max(q0, pmax + 1), b1 = min(q1, hmin);
      if (b0 <= b1) {
        if (q0 < b0) walk(q0, b0 - 1, true);
        walk(b0, b1, false);
        if (b1 < q1) walk(b1 + 1, q1, true);
      } else {
        walk(q0, q1, true);
      }
    }
    if (MODE != kFill && lane == 0) a.counts[it] = em.cnt;
    k = wc ? claim_take(ws, pend) : k + ws.nwaves;
  }
}

// ---- reduced-precision sweep with FP64 confirmation ----
// The 8-byte level records of q (QRec) are tested by two packed 16-bit
// subtractions and one sign-bit test:
//   (hi_p - lo_q) | (hi_q - lo_p) has no negative half  <=>  levels overlap,
// with the kind codes of row p and column q folded in (make_qrec), so kind
// pairs that the type filter drops fail here already.
// A group's rows are close on axis 1 (the planner groups them by group_key),
// so the union of their level boxes is small: positions go across the lanes
// 64 at a time and are tested against the union first (one test for 64
// positions); only the positions that pass are tested against the rows, one
// position at a time, its record read out of the lane that loaded it.
// The (row, q) candidates of a position are written straight into the
// wave's LDS slots (the row lanes of the mask, each at its rank in it, under
// that mask as the execution mask).  Whenever 64 or more wait, 64 of them are
// confirmed at once, one per lane, each lane fetching both exact records:
// exact FP64 overlap on all three axes, type filter, shared-vertex filter.
// One load latency is paid per round of 64 candidates, and the emission order
// is the order of q, then of row lanes, as in the plain sweep.
typedef short short2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bool level_overlap(uint32_t phi, uint32_t plo, uint32_t qlo,
                                              uint32_t qhi) {
  const short2v d1 = __builtin_bit_cast(short2v, phi) - __builtin_bit_cast(short2v, qlo);
  const short2v d2 = __builtin_bit_cast(short2v, qhi) - __builtin_bit_cast(short2v, plo);
  return ((__builtin_bit_cast(uint32_t, d1) | __builtin_bit_cast(uint32_t, d2)) & 0x80008000u) == 0;
}

// Tail items of the reduced-precision sweep: the level test per (row, q), then
// every candidate confirmed at once against both exact records.
template <int MODE>
__device__ __forceinline__ void tailq(const SweepArgs &a, const Exact64 *__restrict__ rec,
                                      const QRec *__restrict__ qrec, int2 *__restrict__ out,
                                      const Item &item, int lane, Emit<MODE> &em,
                                      unsigned long long &nconf) {
  const int R = item.nrows & kRowsMask;
  int rp = INT_MAX, rh = INT_MIN;
  uint32_t rlo = 0x7fff7fffu, rhi = 0u, rop = 0;
  if (lane < R) {
    const int2 r = a.rowlist[item.rowoff + lane];
    rp = r.x;
    rh = r.y;
    const QRec x = row_form(qrec[rp]);
    rlo = x.lo;
    rhi = x.hi;
    if (MODE != kCount) rop = a.orig[rp];
  }
  int r0 = 0;
  for (int qb = item.q0; qb <= item.q1; qb += kWave) {
    while (r0 < R && readlane_i(rh, r0) < qb) ++r0;
    const int q = qb + lane;
    const bool inq = q <= item.q1;
    uint32_t qlo = 0x7fff7fffu, qhi = 0u;
    if (inq) {
      const QRec x = qrec[q];
      qlo = x.lo;
      qhi = x.hi;
    }
    for (int r = r0; r < R; ++r) {
      const int p = readlane_i(rp, r), hp = readlane_i(rh, r);
      const uint32_t plo = (uint32_t)readlane_i((int)rlo, r), phi = (uint32_t)readlane_i((int)rhi, r);
      const bool c = inq & (q > p) & (q <= hp) & level_overlap(phi, plo, qlo, qhi);
      if (ballot(c)) {
        const Exact64 xp = rec[p];
        bool hg = false, h = false;
        if (c) {
          const Exact64 xq = rec[q];
          hg = ((accept_bits(xp.type) & xq.type) != 0) & (xp.hi0 >= xq.lo0) & (xp.lo0 <= xq.hi0) &
               (xp.hi1 >= xq.lo1) & (xp.lo1 <= xq.hi1) & (xp.hi2 >= xq.lo2) & (xp.lo2 <= xq.hi2);
          h = hg & (xp.v0 != xq.v0) & (xp.v0 != xq.v1) & (xp.v0 != xq.v2) & (xp.v1 != xq.v0) &
              (xp.v1 != xq.v1) & (xp.v1 != xq.v2) & (xp.v2 != xq.v0) & (xp.v2 != xq.v1) &
              (xp.v2 != xq.v2);
        }
        nconf += (unsigned long long)__popcll(ballot(hg));
        const uint64_t hm = ballot(h);
        if (hm)
          em.put(h, hm, (uint32_t)readlane_i((int)rop, r), MODE != kCount && h ? a.orig[q] : 0u,
                 out);
      }
    }
  }
}

typedef unsigned short ushort2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_min_u16(uint32_t x, uint32_t y) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(ushort2v, x),
                                                                __builtin_bit_cast(ushort2v, y)));
}
__device__ __forceinline__ uint32_t pk_max_u16(uint32_t x, uint32_t y) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(ushort2v, x),
                                                                __builtin_bit_cast(ushort2v, y)));
}

// The rows of the current item as the confirmation reads them, staged in the
// wave's LDS when the item starts, in four 16-byte planes of 64 rows: the
// FP64 bounds of each axis (sweep axis first), then the vertex ids and the
// input id.  A lane reads one 16-byte entry per plane at 16 * row bytes, so
// 64 lanes reading distinct rows spread evenly over the banks (a 64-byte row
// stride would put every row of the same parity on the same banks).
struct alignas(16) RowBounds {
  double lo, hi;
};
struct alignas(16) RowIds {
  int v0, v1, v2;
  uint32_t orig;
};
struct RowPlanes {
  RowBounds b[3][kWave];
  RowIds id[kWave];
};

// At most 64 VGPRs (the second bound: 8 waves per SIMD), so 32 waves fit on
// a CU: with 71 registers 7 fit, and the learned pass took 7.25 instead of
// 6.81 ms on the 512-scene batch.
template <int MODE>
__global__ __launch_bounds__(kThreads, 8) void k_qsweep(SweepArgs a,
                                                        const Exact64 *__restrict__ rec,
                                                        const QRec *__restrict__ qrec,
                                                        int2 *__restrict__ out) {
  // LDS slots per wave: fewer than 64 waiting candidates plus one position's
  // 64, as (row lane, position); and the item's rows
  __shared__ int2 slot_mem[kThreads / kWave][2 * kWave];
  __shared__ RowPlanes row_mem[kThreads / kWave];
  const int lane = threadIdx.x & 63;
  int2 *slots = slot_mem[threadIdx.x >> 6];
  RowPlanes &rows = row_mem[threadIdx.x >> 6];
  unsigned long long nconf = 0;
  if (MODE == kFill && a.total_dev && *a.total_dev > a.capacity) return;
  const WaveSlice ws = wave_slice(a.nitems_dev ? *a.nitems_dev : a.nitems, a.xcd_split, a.slices);
  // the first item is the wave's own; the claims number the rest from nwaves
  unsigned *wc = a.work ? a.work + ws.xcd * kClaimStride : nullptr;
  int64_t k = ws.kbeg + ws.wave;
  while (k < ws.kend) {
    const unsigned pend = wc ? claim_issue(wc) : 0u;
    const int64_t it = a.sel ? (int64_t)a.sel[k] : k;
    const Item item = a.items[it];
    Emit<MODE> em;
    if (MODE != kCount) em.base0 = em.base = a.offsets[it];
    if (MODE == kLearned) em.cap = a.caps[it];
    if (item.nrows & kTailFlag) {
      tailq<MODE>(a, rec, qrec, out, item, lane, em, nconf);
    } else {
      const bool act = lane < item.nrows;
      const int2 row = act ? a.rowlist[item.rowoff + lane] : make_int2(INT_MAX, INT_MIN);
      // the chain end within the item: the block loads below read past q1
      // unchecked, into positions of other items (the next chunk of a body,
      // the group's tail) or the padding
      const int p = row.x, myhi = min(row.y, item.q1);
      // p < q <= myhi as one unsigned compare: q - (p + 1) < myhi - p
      const int pp1 = p + 1;
      const uint32_t span = act && myhi > p ? (uint32_t)(myhi - p) : 0u;
      QRec mq;
      mq.lo = 0x7fff7fffu;  // idle lanes: an empty level box that meets no column (make_qrec)
      mq.hi = 0u;
      if (act) {
        mq = row_form(qrec[p]);
        const Exact64 x = rec[p];
        rows.b[0][lane] = RowBounds{x.lo0, x.hi0};
        rows.b[1][lane] = RowBounds{x.lo1, x.hi1};
        rows.b[2][lane] = RowBounds{x.lo2, x.hi2};
        rows.id[lane] = RowIds{x.v0, x.v1, x.v2, MODE != kCount ? a.orig[p] : 0u};
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // The union of the rows' level boxes (kind codes dropped): a position
      // whose box misses it meets no row.
      uint32_t ulo = act ? (mq.lo & ~kCodeBits) : 0x3fff3fffu;
      uint32_t uhi = act ? (mq.hi & ~kCodeBits) : 0u;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        ulo = pk_min_u16(ulo, (uint32_t)__shfl_xor((int)ulo, o, kWave));
        uhi = pk_max_u16(uhi, (uint32_t)__shfl_xor((int)uhi, o, kWave));
      }
      ulo = (uint32_t)uniform((int)ulo);
      uhi = (uint32_t)uniform((int)uhi);

      // Candidates wait in this wave's LDS slots as (row lane, q), in
      // emission order (position, then row lane).  Whenever 64 or more wait,
      // 64 are confirmed at once, one per lane: the row's record from the
      // LDS, the position's exact record from memory; the rest (fewer than 64)
      // move to the front.  The kind codes of the level test have already
      // applied the type filter exactly, so the confirmation tests the FP64
      // bounds and the shared vertices.
      int T = 0;
      auto test_slots = [&](int base, int n) __attribute__((always_inline)) {
        bool h = false, hg = false;
        uint32_t o = 0, oq = 0;
        if (lane < n) {
          const int2 c = slots[base + lane];
          // the input id loads with the record, not after the test
          if (MODE != kCount) oq = a.orig[c.y];
          const Exact64 xq = rec[c.y];
          const RowBounds b0 = rows.b[0][c.x], b1 = rows.b[1][c.x], b2 = rows.b[2][c.x];
          const RowIds xp = rows.id[c.x];
          o = xp.orig;
          hg = (b0.hi >= xq.lo0) & (b0.lo <= xq.hi0) & (b1.hi >= xq.lo1) & (b1.lo <= xq.hi1) &
               (b2.hi >= xq.lo2) & (b2.lo <= xq.hi2);
          h = hg;
          h = h & (xp.v0 != xq.v0) & (xp.v0 != xq.v1) & (xp.v0 != xq.v2) & (xp.v1 != xq.v0) &
              (xp.v1 != xq.v1) & (xp.v1 != xq.v2) & (xp.v2 != xq.v0) & (xp.v2 != xq.v1) &
              (xp.v2 != xq.v2);
          // The id is needed only for a pair, and the compiler would load it
          // inside the emission branch, a second memory latency per round
          // after the record's; taking its value here keeps its load issued
          // beside the record's.
          if (MODE != kCount) asm volatile("" : "+v"(oq));
        }
        nconf += (unsigned long long)__popcll(ballot(hg));
        const uint64_t hm = ballot(h);
        if (hm) em.put(h, hm, o, oq, out);
      };
      auto drain_full = [&]() __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        test_slots(0, kWave);
        __builtin_amdgcn_wave_barrier();
        const bool mv = lane < T - kWave;
        int2 v = make_int2(0, 0);
        if (mv) v = slots[kWave + lane];
        __builtin_amdgcn_wave_barrier();
        if (mv) slots[lane] = v;
        __builtin_amdgcn_wave_barrier();
        T -= kWave;
      };
      // Positions go across the lanes, 64 at a time, the next 64 loaded ahead:
      // each lane tests its position against the union; for each position that
      // passes, its record is read out of its lane and tested against every
      // row, and the rows that pass and hold the position in their chain
      // (p < q <= hi) write (p, q) into the next slots, each at its rank.
      const int q0 = item.q0, q1 = item.q1;
      // Every block loads whole: qrec holds kQPad records past its last
      // position, and a position past q1 (another item's, or padding) fails
      // every row's chain test (q <= myhi), whatever its levels.
      // One block: its positions against the union, then each position that
      // passes against every row.
      auto block = [&](int qc, uint32_t xlo, uint32_t xhi) __attribute__((always_inline)) {
        uint64_t pm = ballot(level_overlap(uhi, ulo, xlo & ~kCodeBits, xhi & ~kCodeBits));
        while (pm) {
          const int j = (int)__builtin_ctzll(pm);
          asm("s_bitset0_b64 %0, %1" : "+s"(pm) : "s"(j));  // clears bit j
          const uint32_t lo = (uint32_t)readlane_i((int)xlo, j);
          const uint32_t hi = (uint32_t)readlane_i((int)xhi, j);
          uint64_t m = ballot(level_overlap(mq.hi, mq.lo, lo, hi));
          if (m) {
            const int q = qc + j;
            m &= ballot((uint32_t)(q - pp1) < span);
            if (m) {
              if (__builtin_amdgcn_inverse_ballot_w64(m))
                slots[__builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, (uint32_t)T))] =
                    make_int2(lane, q);
              T += __popcll(m);
              if (T >= kWave) drain_full();
            }
          }
        }
      };
      // Blocks go in pairs, each block's successor loading while it is tested
      // (issued after the wait for its own records, which waits for every
      // load in flight); in pairs, a loaded block's registers are free when
      // the load after next arrives, so no copies are made.
      QRec xa = qrec[q0 + lane];
      const QRec *qn = qrec + ((int64_t)q0 + kWave + lane);  // this lane's next position
      for (int qc = q0; qc <= q1; qc += 2 * kWave, qn += 2 * kWave) {
        const QRec xb = qn[0];
        block(qc, xa.lo, xa.hi);
        if (qc + kWave > q1) break;
        xa = qn[kWave];
        block(qc + kWave, xb.lo, xb.hi);
      }
      if (T) {
        __builtin_amdgcn_wave_barrier();
        test_slots(0, T);
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (MODE != kFill && lane == 0) a.counts[it] = em.cnt;
    k = wc ? claim_take(ws, pend) : k + ws.nwaves;
  }
  if (MODE != kFill && lane == 0 && nconf && a.confirms)
    atomicAdd(a.confirms + ws.xcd * (kClaimStride / 2), nconf);
}

// ---- grouped FP64 sweep ----
// The reduced-precision sweep's organisation with every test in FP64: the
// planner groups rows close on axis 1 (levels only order the rows), positions
// go across the lanes 64 at a time and are tested against the FP64 union box
// of the group's rows on axes 1 and 2, and each position that passes is read
// out of its lane and tested against every row in FP64 (both axes, type,
// shared vertices, p < q <= hi).  Tail items as in the plain FP64 sweep.
// The same 64-VGPR bound as k_qsweep (32 waves per CU).
// Emission order: position, then row lane.
__device__ __forceinline__ double wave_min_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

template <int MODE>
__global__ __launch_bounds__(kThreads, 8) void k_gsweep64(SweepArgs a,
                                                          const Rec64 *__restrict__ rec,
                                                          int2 *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  if (MODE == kFill && a.total_dev && *a.total_dev > a.capacity) return;
  const WaveSlice ws = wave_slice(a.nitems_dev ? *a.nitems_dev : a.nitems, a.xcd_split, a.slices);
  unsigned *wc = a.work ? a.work + ws.xcd * kClaimStride : nullptr;
  int64_t k = ws.kbeg + ws.wave;
  while (k < ws.kend) {
    const unsigned pend = wc ? claim_issue(wc) : 0u;
    const int64_t it = a.sel ? (int64_t)a.sel[k] : k;
    const Item item = a.items[it];
    Emit<MODE> em;
    if (MODE != kCount) em.base0 = em.base = a.offsets[it];
    if (MODE == kLearned) em.cap = a.caps[it];
    if (item.nrows & kTailFlag) {
      tail64<MODE>(a, rec, out, item, lane, em);
    } else {
      const bool act = lane < item.nrows;
      const int2 row = act ? a.rowlist[item.rowoff + lane] : make_int2(INT_MAX, INT_MIN);
      const int p = row.x, myhi = row.y;
      Rec64 mine;
      mine.lo1 = mine.lo2 = INFINITY;  // idle lanes: an empty box
      mine.hi1 = mine.hi2 = -INFINITY;
      mine.v0 = mine.v1 = mine.v2 = mine.type = 0;
      uint32_t op = 0;
      if (act) {
        mine = rec[p];
        if (MODE != kCount) op = a.orig[p];
      }
      const int acc = act ? accept_bits(mine.type) : 0;
      // the union of the rows' boxes on axes 1 and 2
      const double ulo1 = wave_min_f64(mine.lo1), uhi1 = wave_max_f64(mine.hi1);
      const double ulo2 = wave_min_f64(mine.lo2), uhi2 = wave_max_f64(mine.hi2);
      const int q0 = item.q0, q1 = item.q1;
      Rec64 x;
      x.lo1 = x.lo2 = INFINITY;
      x.hi1 = x.hi2 = -INFINITY;
      x.v0 = x.v1 = x.v2 = x.type = 0;
      if (q0 + lane <= q1) x = rec[q0 + lane];
      // Lanes past q1 hold the empty box of type 0, which no row accepts, so
      // the union test needs no range check of its own.
      const Rec64 *xn = rec + ((int64_t)q0 + kWave + lane);  // this lane's next position
      for (int qc = q0; qc <= q1; qc += kWave, xn += kWave) {
        uint64_t pm = ballot((x.lo1 <= uhi1) & (x.hi1 >= ulo1) & (x.lo2 <= uhi2) & (x.hi2 >= ulo2));
        // the next 64 positions load while these are tested; all but an
        // item's last block are whole, and load without a per-lane check
        Rec64 nx;
        nx.lo1 = nx.lo2 = INFINITY;
        nx.hi1 = nx.hi2 = -INFINITY;
        nx.v0 = nx.v1 = nx.v2 = nx.type = 0;
        if (qc + 2 * kWave - 1 <= q1)
          nx = *xn;
        else if (qc + kWave + lane <= q1)
          nx = *xn;
        uint32_t oql = 0;
        if (MODE != kCount && pm && qc + lane <= q1) oql = a.orig[qc + lane];
        while (pm) {
          const int j = (int)__builtin_ctzll(pm);
          pm &= pm - 1;
          const int q = qc + j;
          const double lo1 = readlane_f64(x.lo1, j), hi1 = readlane_f64(x.hi1, j);
          const double lo2 = readlane_f64(x.lo2, j), hi2 = readlane_f64(x.hi2, j);
          const int qt = readlane_i(x.type, j);
          const bool g = (q > p) & (q <= myhi) & (mine.lo1 <= hi1) & (mine.hi1 >= lo1) &
                         (mine.lo2 <= hi2) & (mine.hi2 >= lo2) & ((acc & qt) != 0);
          if (ballot(g)) {
            const int v0 = readlane_i(x.v0, j), v1 = readlane_i(x.v1, j), v2 = readlane_i(x.v2, j);
            const bool h = g & (mine.v0 != v0) & (mine.v0 != v1) & (mine.v0 != v2) &
                           (mine.v1 != v0) & (mine.v1 != v1) & (mine.v1 != v2) & (mine.v2 != v0) &
                           (mine.v2 != v1) & (mine.v2 != v2);
            const uint64_t hm = ballot(h);
            if (hm) em.put(h, hm, op, MODE != kCount ? (uint32_t)readlane_i((int)oql, j) : 0u, out);
          }
        }
        x = nx;
      }
    }
    if (MODE != kFill && lane == 0) a.counts[it] = em.cnt;
    k = wc ? claim_take(ws, pend) : k + ws.nwaves;
  }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

// ---- host wrappers -------------------------------------------------------------------
void launch_axis_stats(const stq_aabb_f64 *boxes, int64_t n, double *partial, int nblocks,
                       hipStream_t st, uint32_t *keys3, unsigned *err) {
  hipLaunchKernelGGL(k_axis_stats, dim3(nblocks), dim3(kThreads), 0, st, boxes, n, partial, keys3,
                     err);
  STQ_HIP(hipGetLastError());
}

void launch_axis_final(const double *partial, int nblocks, int64_t n, int policy,
                       SceneParams *out, hipStream_t st) {
  hipLaunchKernelGGL(k_axis_final, dim3(1), dim3(kThreads), 0, st, partial, nblocks, n, policy,
                     sweep_key_bits(n), out);
  STQ_HIP(hipGetLastError());
}

void launch_sweep_range(const stq_aabb_f64 *boxes, const int64_t *seg_off, int nscenes, int64_t n,
                        SceneParams *sp, hipStream_t st) {
  if (nscenes <= 0) return;
  hipLaunchKernelGGL(k_sweep_range, dim3((unsigned)nscenes), dim3(kThreads), 0, st, boxes, seg_off,
                     n, sp);
  STQ_HIP(hipGetLastError());
}

void launch_keys(bool f32, const stq_aabb_f64 *boxes, int64_t n, int axis, const SceneParams *sp,
                 void *keys, uint32_t *vals, unsigned *err, hipStream_t st) {
  if (n <= 0) return;
  if (f32)
    hipLaunchKernelGGL(k_keys_f32, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, n, axis, sp,
                       (uint32_t *)keys, vals, err);
  else
    hipLaunchKernelGGL(k_keys_f64, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, n, axis, sp,
                       (uint64_t *)keys, vals, err);
  STQ_HIP(hipGetLastError());
}

void launch_gather(bool f32, const stq_aabb_f64 *boxes, const uint32_t *order, int64_t n,
                   int axis, const SceneParams *sp, void *smin, void *smax, void *rec, QRec *qrec,
                   ColStats *cols, uint2 *rowf, hipStream_t st) {
  if (n <= 0) return;
  if (f32)
    hipLaunchKernelGGL(k_gather_f32, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, order,
                       n, sp, (uint32_t *)smin, (uint32_t *)smax, (Exact64 *)rec, qrec, cols, rowf);
  else
    hipLaunchKernelGGL(k_gather_f64, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, order,
                       n, axis, sp, (double *)smin, (double *)smax, (Rec64 *)rec, qrec, cols, rowf);
  STQ_HIP(hipGetLastError());
}

void launch_chain_end(bool f32, const void *smin, const void *smax, int64_t n, int *hi,
                      unsigned long long *cand, hipStream_t st, int64_t r0, int64_t r1) {
  if (n <= 0) return;
  if (r1 < 0 || r1 > n) r1 = n;
  // two blocks per CU, persistent over the tiles (one counter atomic per block)
  const int64_t rows = f32 ? kCeRows<uint32_t> : kCeRows<double>;
  const unsigned g = (unsigned)std::min<int64_t>((n + rows - 1) / rows, 512);
  if (f32)
    hipLaunchKernelGGL(k_chain_end<uint32_t>, dim3(g), dim3(kCeThreads), 0, st,
                       (const uint32_t *)smin, (const uint32_t *)smax, n, r0, r1, hi, cand);
  else
    hipLaunchKernelGGL(k_chain_end<double>, dim3(g), dim3(kCeThreads), 0, st, (const double *)smin,
                       (const double *)smax, n, r0, r1, hi, cand);
  STQ_HIP(hipGetLastError());
}

void launch_plan(const int *hi, const QRec *qrec, int64_t n, int W, int chunk, float tail_cost,
                 int2 *rowlist, int4 *groups, int4 *tgroups, int2 *gchunk, uint32_t *nchunks,
                 unsigned long long *lane_steps, hipStream_t st, bool cap, const uint2 *rowf,
                 uint2 *rowfeat) {
  if (n <= 0) return;
  if (W > (qrec ? kMaxWindow32 : kMaxWindow64))
    throw StatusError(STQ_ERR_INVALID_ARGUMENT, "planning window too large");
  const int64_t nwin = (n + W - 1) / W;
  const dim3 grid((unsigned)std::min<int64_t>(nwin, kReduceGrid));
  if (qrec)
    hipLaunchKernelGGL(k_plan<true>, grid, dim3(kPlanThreads), 0, st, hi, qrec, n, W, chunk,
                       tail_cost, cap, rowlist, groups, tgroups, gchunk, nchunks, lane_steps, rowf,
                       rowfeat);
  else
    hipLaunchKernelGGL(k_plan<false>, grid, dim3(kPlanThreads), 0, st, hi, qrec, n, W, chunk,
                       tail_cost, cap, rowlist, groups, tgroups, gchunk, nchunks, lane_steps, rowf,
                       rowfeat);
  STQ_HIP(hipGetLastError());
}

void launch_items(const int4 *groups, const int4 *tgroups, const int2 *gchunk, const int64_t *ioff,
                  int64_t ngroups, Item *items, hipStream_t st) {
  if (ngroups <= 0) return;
  hipLaunchKernelGGL(k_items, dim3(blocks_for(ngroups)), dim3(kThreads), 0, st, groups, tgroups,
                     gchunk, ioff, ngroups, items);
  STQ_HIP(hipGetLastError());
}

void launch_item_cost(const Item *items, int64_t n, const int64_t *n_dev, bool grouped,
                      const float *pred, const uint32_t *cand, const uint32_t *counts,
                      uint32_t *cost, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_item_cost, dim3(blocks_for(n)), dim3(kThreads), 0, st, items, n, n_dev,
                     grouped, pred, cand, counts, cost);
  STQ_HIP(hipGetLastError());
}

void launch_xcd_slices(const int64_t *prefix, int64_t n, const int64_t *n_dev, int64_t *slices,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_xcd_slices, dim3(1), dim3(64), 0, st, prefix, n, n_dev, slices);
  STQ_HIP(hipGetLastError());
}

template <int B>
static void launch_sweep_b(bool f32, int mode, const dim3 &grid, const SweepArgs &args,
                           hipStream_t st) {
  const dim3 block(kThreads);
  if (f32) {
    const Exact64 *r = static_cast<const Exact64 *>(args.rec);
    if (mode == kCount)
      hipLaunchKernelGGL((k_qsweep<kCount>), grid, block, 0, st, args, r, args.qrec, args.out);
    else if (mode == kFill)
      hipLaunchKernelGGL((k_qsweep<kFill>), grid, block, 0, st, args, r, args.qrec, args.out);
    else
      hipLaunchKernelGGL((k_qsweep<kLearned>), grid, block, 0, st, args, r, args.qrec, args.out);
  } else if (args.qrec) {  // grouped FP64
    const Rec64 *r = static_cast<const Rec64 *>(args.rec);
    if (mode == kCount)
      hipLaunchKernelGGL((k_gsweep64<kCount>), grid, block, 0, st, args, r, args.out);
    else if (mode == kFill)
      hipLaunchKernelGGL((k_gsweep64<kFill>), grid, block, 0, st, args, r, args.out);
    else
      hipLaunchKernelGGL((k_gsweep64<kLearned>), grid, block, 0, st, args, r, args.out);
  } else {
    constexpr int B64 = B > 4 ? 4 : B;  // Rec64 batches above 4 exceed the SGPR file
    const Rec64 *r = static_cast<const Rec64 *>(args.rec);
    if (mode == kCount)
      hipLaunchKernelGGL((k_sweep64<kCount, B64>), grid, block, 0, st, args, r, args.out);
    else if (mode == kFill)
      hipLaunchKernelGGL((k_sweep64<kFill, B64>), grid, block, 0, st, args, r, args.out);
    else
      hipLaunchKernelGGL((k_sweep64<kLearned, B64>), grid, block, 0, st, args, r, args.out);
  }
}

void launch_sweep(bool f32, int mode, int batch, int nblocks, const SweepArgs &args,
                  hipStream_t st) {
  if (args.nitems <= 0) return;
  // XCD slices: every XCD gets gridDim / kXcds blocks, and each wave's first
  // item is its index in the slice, so a remainder block would repeat items
  if (nblocks <= 0 || (args.xcd_split && nblocks % kXcds != 0))
    throw StatusError(STQ_ERR_INVALID_ARGUMENT, "sweep grid: a positive multiple of 8 blocks");
  const dim3 grid(nblocks);
  // the batch of scalar-loaded positions applies to the FP64 sweep only
  if (f32 || batch <= 2) launch_sweep_b<2>(f32, mode, grid, args, st);
  else launch_sweep_b<4>(f32, mode, grid, args, st);
  STQ_HIP(hipGetLastError());
}


}  // namespace stqamd
