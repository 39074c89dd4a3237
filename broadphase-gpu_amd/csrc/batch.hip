// batch.hip -- kernels of the batched pipeline (see batch.hpp).
#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

#include "batch.hpp"
#include "plan.hpp"

namespace stqamd {

namespace {

constexpr int kT = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

__device__ inline void other_axes_b(int axis, int &a1, int &a2) {
  a1 = axis == 0 ? 1 : 0;
  a2 = axis == 2 ? 1 : 2;
}

// ---- per-scene statistics ---------------------------------------------------
// With keys3 (reduced precision) the same pass writes the sort keys of every
// axis, keys3[a n + i] = image of the rounded-down minimum on axis a, and
// flags NaN or inverted boxes: the sort then reads its scene's axis, so the
// boxes are read once before the sort instead of twice.
__global__ __launch_bounds__(kT) void k_seg_stats(const stq_aabb_f64 *__restrict__ boxes,
                                                  const int64_t *__restrict__ cstart,
                                                  const int64_t *__restrict__ cend,
                                                  double *__restrict__ partial,
                                                  uint32_t *__restrict__ keys3, int64_t n,
                                                  unsigned *__restrict__ err) {
  const int64_t b0 = cstart[blockIdx.x], b1 = cend[blockIdx.x];
  double acc[12], lo[3], up[3];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = INFINITY;
    up[a] = -INFINITY;
  }
  if (!keys3) {
    axis_stats_range(boxes, b0 + threadIdx.x, b1, kT, acc, lo, up);
  } else {
    bool bad = false;
    axis_stats_keys_range(boxes, b0 + threadIdx.x, b1, kT, acc, lo, up, keys3, n, bad);
    if (bad) atomicOr(err, kErrInverted);
  }
  block_axis_stats<kT>(acc, lo, up, partial + (int64_t)blockIdx.x * kAxisStats);
}

// One thread per scene: its chunks are combined in order, then the axis is
// chosen (policy as in the single-scene path) and the level range set.
__global__ __launch_bounds__(kT) void k_seg_finalize(const double *__restrict__ partial,
                                                     const int *__restrict__ scb,
                                                     const int64_t *__restrict__ seg_off,
                                                     int nscenes, int policy, int fixed_axis,
                                                     int kbits, SceneParams *__restrict__ out) {
  const int s = blockIdx.x * kT + threadIdx.x;
  if (s >= nscenes) return;
  double v[kAxisStats];
  for (int k = 0; k < kAxisStats; ++k) v[k] = k < 12 ? 0.0 : (k < 15 ? INFINITY : -INFINITY);
  for (int c = scb[s]; c < scb[s + 1]; ++c)
    for (int k = 0; k < kAxisStats; ++k) {
      const double x = partial[(int64_t)c * kAxisStats + k];
      v[k] = k < 12 ? v[k] + x : (k < 15 ? fmin(v[k], x) : fmax(v[k], x));
    }
  SceneParams p;
  for (int a = 0; a < 3; ++a) {
    const double lo = v[12 + a], up = v[15 + a], range = up - lo;
    p.qp.origin[a] = isfinite(lo) ? lo : 0.0;
    double sc = (range > 0 && isfinite(range)) ? kLevelMax / range : 0.0;
    p.qp.scale[a] = isfinite(sc) ? sc : 0.0;
  }
  const double dn = (double)(seg_off[s + 1] - seg_off[s]);
  int best = fixed_axis;
  if (policy >= 3 && dn > 0) {
    double bestv = 0;
    for (int a = 0; a < 3; ++a) {
      const double s1 = v[4 * a], s2 = v[4 * a + 1], ext = v[4 * a + 2], cc = v[4 * a + 3];
      const double mu = s1 / (2.0 * dn);
      if (policy == 3) {
        const double var = s2 - 2.0 * mu * s1 + 2.0 * dn * mu * mu;
        if (a == 0 || var > bestv) { best = a; bestv = var; }
      } else {
        const double sd = sqrt(fmax(cc / dn - mu * mu, 0.0));
        const double score = sd > 0 ? (ext / dn) / sd : INFINITY;
        if (a == 0 || score < bestv) { best = a; bestv = score; }
      }
    }
  }
  p.axis = best;
  set_sweep_keys(p, v[12 + best], v[15 + best], kbits);
  out[s] = p;
}

// ---- keys -------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_keys_seg(const stq_aabb_f64 *__restrict__ boxes, int64_t n,
                                                 const int64_t *__restrict__ seg_off, int nscenes,
                                                 const SceneParams *__restrict__ sp, int f32,
                                                 void *__restrict__ keys,
                                                 uint32_t *__restrict__ vals,
                                                 unsigned *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  __shared__ int sc0;  // the scene of the block's first box, found once
  if (threadIdx.x == 0) sc0 = scene_of(seg_off, nscenes, (int64_t)blockIdx.x * kT);
  __syncthreads();
  if (i >= n) return;
  int s = sc0;
  while (seg_off[s + 1] <= i) ++s;
  const int axis = sp[s].axis;
  const stq_aabb_f64 &b = boxes[i];
  const double x0 = b.min[0], x1 = b.min[1], x2 = b.min[2];
  const double y0 = b.max[0], y1 = b.max[1], y2 = b.max[2];
  if (!(x0 <= y0) || !(x1 <= y1) || !(x2 <= y2)) atomicOr(err, kErrInverted);
  const double m = axis == 0 ? x0 : (axis == 1 ? x1 : x2);
  if (f32) static_cast<uint32_t *>(keys)[i] = key_of(round_down_f32(m));
  else static_cast<uint64_t *>(keys)[i] = key_of(m);
  vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kT) void k_scene_keys(const uint32_t *__restrict__ vals, int64_t n,
                                                   const int64_t *__restrict__ seg_off,
                                                   int nscenes, uint32_t *__restrict__ sk) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) sk[i] = (uint32_t)scene_of(seg_off, nscenes, (int64_t)vals[i]);
}

// ---- gather --------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_gather_seg(const stq_aabb_f64 *__restrict__ boxes,
                                                   const uint32_t *__restrict__ order, int64_t n,
                                                   const int64_t *__restrict__ seg_off,
                                                   int nscenes,
                                                   const SceneParams *__restrict__ sp, int f32,
                                                   void *__restrict__ smin, void *__restrict__ smax,
                                                   void *__restrict__ rec, QRec *__restrict__ qrec,
                                                   uint32_t *__restrict__ orig_local,
                                                   ColStats *__restrict__ cols,
                                                   uint2 *__restrict__ rowf) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  const bool valid = s < n;
  // the scene of the block's first position, found once; a block of kT
  // consecutive positions meets at most a few scenes after it
  __shared__ int sc0;
  if (threadIdx.x == 0) sc0 = scene_of(seg_off, nscenes, (int64_t)blockIdx.x * kT);
  __syncthreads();
  double lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0, o1 = 0, n1 = 0, o2 = 0, n2 = 0;
  int type = 0;
  if (valid) {
    int sc = sc0;
    while (seg_off[sc + 1] <= s) ++sc;
    const uint32_t o = order[s];
    orig_local[s] = (uint32_t)((int64_t)o - seg_off[sc]);
    const stq_aabb_f64 &b = boxes[o];
    const SceneParams p = sp[sc];
    const int axis = p.axis;
    int a1, a2;
    other_axes_b(axis, a1, a2);
    lo1 = b.min[a1];
    hi1 = b.max[a1];
    lo2 = b.min[a2];
    hi2 = b.max[a2];
    type = type_bit(b.vertex_ids[1], b.vertex_ids[2]);
    o1 = p.qp.origin[a1];
    n1 = p.qp.scale[a1] / kLevelMax;
    o2 = p.qp.origin[a2];
    n2 = p.qp.scale[a2] / kLevelMax;
    if (f32) {
      static_cast<uint32_t *>(smin)[s] = skey_down(round_down_f32(b.min[axis]), p);  // sort key
      static_cast<uint32_t *>(smax)[s] = skey_up(round_up_f32(b.max[axis]), p);
      Exact64 r;
      r.lo0 = b.min[axis];
      r.hi0 = b.max[axis];
      r.lo1 = lo1;
      r.hi1 = hi1;
      r.lo2 = lo2;
      r.hi2 = hi2;
      r.v0 = b.vertex_ids[0];
      r.v1 = b.vertex_ids[1];
      r.v2 = b.vertex_ids[2];
      r.type = type;
      static_cast<Exact64 *>(rec)[s] = r;
      qrec[s] = make_qrec(r.lo1, r.hi1, r.lo2, r.hi2, r.type, p.qp.origin[a1], p.qp.scale[a1],
                          p.qp.origin[a2], p.qp.scale[a2]);
    } else {
      static_cast<double *>(smin)[s] = b.min[axis];
      static_cast<double *>(smax)[s] = b.max[axis];
      Rec64 r;
      r.lo1 = lo1;
      r.hi1 = hi1;
      r.lo2 = lo2;
      r.hi2 = hi2;
      r.v0 = b.vertex_ids[0];
      r.v1 = b.vertex_ids[1];
      r.v2 = b.vertex_ids[2];
      r.type = type;
      static_cast<Rec64 *>(rec)[s] = r;
      // levels for the axis-1 grouping of the grouped FP64 form
      if (qrec)
        qrec[s] = make_qrec(r.lo1, r.hi1, r.lo2, r.hi2, r.type, p.qp.origin[a1], p.qp.scale[a1],
                            p.qp.origin[a2], p.qp.scale[a2]);
    }
  }
  if (rowf && valid) rowf[s] = row_feat(lo1, hi1, lo2, hi2, type, o1, n1, o2, n2);
  // blocks that straddle two scenes mix both (the features are estimates)
  if (cols) col_stats_put(valid, s, lo1, hi1, lo2, hi2, type, o1, n1, o2, n2, cols);
}

// ---- chain ends, bounded by the scene ---------------------------------------
// As k_chain_end: tiles of rows with their minima in LDS; a row's search
// stops at the end of its scene.
template <class T>
__global__ __launch_bounds__(kCeThreads, 8) void k_chain_end_seg(const T *__restrict__ smin,
                                                              const T *__restrict__ smax,
                                                              int64_t n,
                                                              const int64_t *__restrict__ seg_off,
                                                              int nscenes, int *__restrict__ hi,
                                                              unsigned long long *__restrict__ cand) {
  __shared__ __align__(16) T win[kCeWin<T>];  // 16-byte window stores
  __shared__ int sc_range[2];
  unsigned long long len = 0;
  const int64_t ntiles = (n + kCeRows<T> - 1) / kCeRows<T>;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t s0 = t * kCeRows<T>;
    const int64_t s1 = s0 + kCeRows<T> < n ? s0 + kCeRows<T> : n;
    if (threadIdx.x == 0) {  // the scenes of the tile's first and last rows
      sc_range[0] = scene_of(seg_off, nscenes, s0);
      sc_range[1] = scene_of(seg_off, nscenes, s1 - 1);
    }
    __syncthreads();
    const int sa = sc_range[0], sb = sc_range[1];
    // the end of a row's scene: one lookup when the tile lies in one scene,
    // else a search among the tile's scenes
    auto end_of = [=](int64_t s) {
      int lo = sa, up = sb + 1;
      while (up - lo > 1) {
        const int mid = (lo + up) >> 1;
        if (seg_off[mid] <= s) lo = mid; else up = mid;
      }
      return seg_off[lo + 1];
    };
    len += chain_end_tile(smin, smax, n, s0, (int64_t)0, n, win, end_of, hi);
  }
  block_atomic_add<kCeThreads>(len, cand);
}

// ---- planning -------------------------------------------------------------------
template <bool Q>
__global__ __launch_bounds__(kPlanThreads) void k_plan_seg(const int *__restrict__ hi,
                                                 const QRec *__restrict__ qrec,
                                                 const int64_t *__restrict__ win_start,
                                                 const int *__restrict__ win_cnt,
                                                 const int64_t *__restrict__ win_gbase, int nwin,
                                                 int W, int chunk, float tail_cost,
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
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long steps = 0;
  for (int b = blockIdx.x; b < nwin; b += gridDim.x) {  // grid-stride over windows
    const int64_t wbase = win_start[b];
    const int cnt = win_cnt[b];
    const int64_t gbase = win_gbase[b];
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
      // capped: the item buffer is sized on the host from the scene sizes
      const GroupPlan gp = plan_group(lane, nrows, p, h, (int)(wbase + 64 * k), chunk, true,
                                      tail_cost);
      const int64_t g = gbase + k;
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

__global__ __launch_bounds__(kT) void k_scene_pairs(const int64_t *__restrict__ off,
                                                    const int64_t *__restrict__ ioff,
                                                    const int64_t *__restrict__ sfg, int nscenes,
                                                    int64_t *__restrict__ pair_off) {
  const int s = blockIdx.x * kT + threadIdx.x;
  if (s <= nscenes) pair_off[s] = off[ioff[sfg[s]]];
}


__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

void build_batch_tables(const int64_t *off, int nscenes, int W, BatchTables &t) {
  t.nscenes = nscenes;
  t.seg_off.assign(off, off + nscenes + 1);
  t.n = off[nscenes];
  t.max_scene = 0;
  for (int s = 0; s < nscenes; ++s) t.max_scene = std::max(t.max_scene, off[s + 1] - off[s]);
  t.chunk_scene.clear();
  t.chunk_start.clear();
  t.chunk_end.clear();
  t.scene_chunk_begin.assign(nscenes + 1, 0);
  t.win_start.clear();
  t.win_cnt.clear();
  t.win_gbase.clear();
  t.scene_first_group.assign(nscenes + 1, 0);
  t.sort_tiles.clear();
  const int tk = sort_tile_keys();
  for (int s = 0; s < nscenes; ++s) {
    const int64_t a = off[s], m = off[s + 1] - off[s];
    const int64_t nt = (m + tk - 1) / tk, t0 = (int64_t)t.sort_tiles.size();
    for (int64_t i = 0; i < nt; ++i) {
      SegTile x;
      x.start = a + i * tk;
      x.count = (int32_t)std::min<int64_t>(tk, m - i * tk);
      x.cbase = 256 * t0 + i;
      x.stride = (int32_t)nt;
      x.scene = s;
      x.pad = 0;
      t.sort_tiles.push_back(x);
    }
  }
  int64_t g = 0;
  const int64_t kStatChunk = 8192;  // 32 boxes per thread of a statistics block
  for (int s = 0; s < nscenes; ++s) {
    const int64_t a = off[s], b = off[s + 1];
    t.scene_chunk_begin[s] = (int)t.chunk_scene.size();
    for (int64_t c = a; c < b; c += kStatChunk) {
      t.chunk_scene.push_back(s);
      t.chunk_start.push_back(c);
      t.chunk_end.push_back(std::min(b, c + kStatChunk));
    }
    t.scene_first_group[s] = g;
    if (b - a > 1) {
      for (int64_t w = a; w < b; w += W) {
        const int cnt = (int)std::min<int64_t>(W, b - w);
        t.win_start.push_back(w);
        t.win_cnt.push_back(cnt);
        t.win_gbase.push_back(g);
        g += (cnt + 63) / 64;
      }
    }
  }
  t.scene_chunk_begin[nscenes] = (int)t.chunk_scene.size();
  t.scene_first_group[nscenes] = g;
  t.ngroups = g;
}

void launch_seg_stats(const stq_aabb_f64 *boxes, const BatchTables &t, const BatchDev &d,
                      int policy, int fixed_axis, uint32_t *keys3, unsigned *err, hipStream_t st) {
  const int nc = (int)t.chunk_scene.size();
  if (nc > 0)
    hipLaunchKernelGGL(k_seg_stats, dim3(nc), dim3(kT), 0, st, boxes, d.chunk_start, d.chunk_end,
                       d.partial, keys3, t.n, err);
  if (t.nscenes > 0)
    hipLaunchKernelGGL(k_seg_finalize, dim3(nblk(t.nscenes)), dim3(kT), 0, st,
                       (const double *)d.partial, d.scene_chunk_begin, d.seg_off, t.nscenes, policy,
                       fixed_axis, sweep_key_bits(t.max_scene), d.params);
  STQ_HIP(hipGetLastError());
  if (keys3) launch_sweep_range(boxes, d.seg_off, t.nscenes, t.n, d.params, st);
}

void launch_keys_seg(bool f32, const stq_aabb_f64 *boxes, const BatchTables &t, const BatchDev &d,
                     void *keys, uint32_t *vals, unsigned *err, hipStream_t st) {
  if (t.n <= 0) return;
  hipLaunchKernelGGL(k_keys_seg, dim3(nblk(t.n)), dim3(kT), 0, st, boxes, t.n, d.seg_off, t.nscenes,
                     (const SceneParams *)d.params, f32 ? 1 : 0, keys, vals, err);
  STQ_HIP(hipGetLastError());
}

void launch_scene_keys(const uint32_t *vals, const BatchTables &t, const BatchDev &d,
                       uint32_t *scene_key, hipStream_t st) {
  if (t.n <= 0) return;
  hipLaunchKernelGGL(k_scene_keys, dim3(nblk(t.n)), dim3(kT), 0, st, vals, t.n, d.seg_off,
                     t.nscenes, scene_key);
  STQ_HIP(hipGetLastError());
}

void launch_gather_seg(bool f32, const stq_aabb_f64 *boxes, const uint32_t *order,
                       const BatchTables &t, const BatchDev &d, void *smin, void *smax, void *rec,
                       QRec *qrec, uint32_t *orig_local, ColStats *cols, uint2 *rowf,
                       hipStream_t st) {
  if (t.n <= 0) return;
  hipLaunchKernelGGL(k_gather_seg, dim3(nblk(t.n)), dim3(kT), 0, st, boxes, order, t.n, d.seg_off,
                     t.nscenes, (const SceneParams *)d.params, f32 ? 1 : 0, smin, smax, rec, qrec,
                     orig_local, cols, rowf);
  STQ_HIP(hipGetLastError());
}

void launch_chain_end_seg(bool f32, const void *smin, const void *smax, const BatchTables &t,
                          const BatchDev &d, int *hi, unsigned long long *cand, hipStream_t st) {
  if (t.n <= 0) return;
  // two blocks per CU, persistent over the tiles (one counter atomic per block)
  const int64_t rows = f32 ? kCeRows<uint32_t> : kCeRows<double>;
  const unsigned g = (unsigned)std::min<int64_t>((t.n + rows - 1) / rows, 512);
  if (f32)
    hipLaunchKernelGGL(k_chain_end_seg<uint32_t>, dim3(g), dim3(kCeThreads), 0, st,
                       (const uint32_t *)smin, (const uint32_t *)smax, t.n, d.seg_off, t.nscenes,
                       hi, cand);
  else
    hipLaunchKernelGGL(k_chain_end_seg<double>, dim3(g), dim3(kCeThreads), 0, st,
                       (const double *)smin, (const double *)smax, t.n, d.seg_off, t.nscenes, hi,
                       cand);
  STQ_HIP(hipGetLastError());
}

void launch_plan_seg(const int *hi, const QRec *qrec, const BatchTables &t, const BatchDev &d,
                     int W, int chunk, float tail_cost, int2 *rowlist, int4 *groups,
                     int4 *tgroups, int2 *gchunk, uint32_t *nchunks,
                     unsigned long long *lane_steps, hipStream_t st, const uint2 *rowf,
                     uint2 *rowfeat) {
  const int nw = (int)t.win_start.size();
  if (nw <= 0) return;
  if (W > (qrec ? kMaxWindow32 : kMaxWindow64))
    throw StatusError(STQ_ERR_INVALID_ARGUMENT, "planning window too large");
  hipLaunchKernelGGL(qrec ? k_plan_seg<true> : k_plan_seg<false>,
                     dim3(std::min((unsigned)nw, kReduceGrid)), dim3(kPlanThreads), 0, st, hi,
                     qrec,
                     d.win_start, d.win_cnt, d.win_gbase, nw, W, chunk, tail_cost, rowlist, groups,
                     tgroups, gchunk, nchunks, lane_steps, rowf, rowfeat);
  STQ_HIP(hipGetLastError());
}

void launch_scene_pairs(const int64_t *off, const int64_t *ioff, const BatchTables &t,
                        const BatchDev &d, int64_t *pair_off, hipStream_t st) {
  hipLaunchKernelGGL(k_scene_pairs, dim3(nblk(t.nscenes + 1)), dim3(kT), 0, st, off, ioff,
                     d.scene_first_group, t.nscenes, pair_off);
  STQ_HIP(hipGetLastError());
}


void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(nblk(n)), dim3(kT), 0, st, p, v, n);
  STQ_HIP(hipGetLastError());
}
void launch_fill_u64(unsigned long long *p, unsigned long long v, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(nblk(n)), dim3(kT), 0, st, p, v, n);
  STQ_HIP(hipGetLastError());
}

}  // namespace stqamd
