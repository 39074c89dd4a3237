// sizer.hip -- features, MLP inference/training and the compaction of the
// learned single-pass sweep (see sizer.hpp).  Small kernels: the network has
// 6,337 parameters and a scene has at most a few hundred thousand items.
#include <climits>
#include <cmath>

#include "plan.hpp"
#include "sizer.hpp"

namespace stqamd {

namespace {

constexpr int kT = 256;
#ifndef STQ_GRAD_WAVES
#define STQ_GRAD_WAVES 4
#endif

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Sum over the 64 lanes of a fully active wave with DPP row moves instead of
// LDS permutes: pairs, quads (quad_perm), 8 and 16 lanes (row half-mirror,
// row mirror), then rows 0 -> 1 and 2 -> 3 (row_bcast:15 into rows 1 and 3)
// and rows 0..1 -> 2..3 (row_bcast:31 into rows 2 and 3); lane 63 then holds
// the total, which is broadcast.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_move(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v += dpp_move<0xB1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4E, 0xf>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xf>(v);  // row_mirror
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// ---- prefix sums of the column statistics ---------------------------------------
// pre[j][k] = sum of field k (ColStats order) over blocks 0 .. j - 1, in
// FP64, j = 0 .. m; an item's column sums are then two lookups.  Three
// launches: the sum of every run of kColRun blocks (one thread block each,
// kColPer consecutive blocks per thread), the exclusive scan of the runs (one
// thread block), each block's prefix (a thread-block scan on top of its run's
// prefix).
constexpr int kColFields = 12;
constexpr int kColPer = 4;
constexpr int kColRun = kT * kColPer;
constexpr int kTopThreads = 1024;

__device__ __forceinline__ void col_fields(const ColStats &c, double *v) {
  v[0] = c.le1;
  v[1] = c.le1sq;
  v[2] = c.le2;
  v[3] = c.le2sq;
  v[4] = c.c1;
  v[5] = c.c1sq;
  v[6] = c.c2;
  v[7] = c.c2sq;
  v[8] = c.nv;
  v[9] = c.ne;
  v[10] = c.nf;
  v[11] = c.cnt;
}

// Sum of the fields over the thread block (fixed order): wave trees, then the
// waves in order; the result is valid in every thread.
__device__ __forceinline__ void block_sum_fields(double *v) {
  __shared__ double wt[kT / 64][kColFields];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kColFields; ++k) {
    v[k] = wave_sum(v[k]);
    if (lane == 0) wt[w][k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kColFields; ++k) {
    double x = 0.0;
    for (int i = 0; i < kT / 64; ++i) x += wt[i][k];
    v[k] = x;
  }
}

__global__ __launch_bounds__(kT) void k_col_runs(const ColStats *__restrict__ cols, int64_t m,
                                                 double *__restrict__ run) {
  const int64_t j0 = (int64_t)blockIdx.x * kColRun + (int64_t)threadIdx.x * kColPer;
  double v[kColFields];
#pragma unroll
  for (int k = 0; k < kColFields; ++k) v[k] = 0.0;
#pragma unroll
  for (int i = 0; i < kColPer; ++i) {
    if (j0 + i < m) {
      double f[kColFields];
      col_fields(cols[j0 + i], f);
#pragma unroll
      for (int k = 0; k < kColFields; ++k) v[k] += f[k];
    }
  }
  block_sum_fields(v);
  if (threadIdx.x < kColFields) {
    double x = v[0];
#pragma unroll
    for (int k = 1; k < kColFields; ++k) x = (int)threadIdx.x == k ? v[k] : x;
    run[(int64_t)blockIdx.x * kColFields + threadIdx.x] = x;
  }
}

// Exclusive scan of the nr run sums in place; the total goes to total[].
// Thread t owns a contiguous slice of runs.
__global__ __launch_bounds__(kTopThreads) void k_col_top(double *__restrict__ run, int64_t nr,
                                                         double *__restrict__ total) {
  __shared__ double wtot[kTopThreads / 64][kColFields];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t per = (nr + kTopThreads - 1) / kTopThreads;
  const int64_t r0 = (int64_t)t * per < nr ? (int64_t)t * per : nr;
  const int64_t r1 = r0 + per < nr ? r0 + per : nr;
  double own[kColFields], inc[kColFields];
#pragma unroll
  for (int k = 0; k < kColFields; ++k) own[k] = 0.0;
  for (int64_t r = r0; r < r1; ++r)
#pragma unroll
    for (int k = 0; k < kColFields; ++k) own[k] += run[r * kColFields + k];
#pragma unroll
  for (int k = 0; k < kColFields; ++k) {
    double x = own[k];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double y = __shfl_up(x, o, kWave);
      if (lane >= o) x += y;
    }
    inc[k] = x;
    if (lane == 63) wtot[w][k] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kColFields; ++k) {
    double before = 0.0;
    for (int i = 0; i < w; ++i) before += wtot[i][k];
    inc[k] += before;
  }
  if (t == kTopThreads - 1) {
#pragma unroll
    for (int k = 0; k < kColFields; ++k) total[k] = inc[k];
  }
  double acc[kColFields];
#pragma unroll
  for (int k = 0; k < kColFields; ++k) acc[k] = inc[k] - own[k];
  for (int64_t r = r0; r < r1; ++r)
#pragma unroll
    for (int k = 0; k < kColFields; ++k) {
      const double x = run[r * kColFields + k];
      run[r * kColFields + k] = acc[k];
      acc[k] += x;
    }
}

__global__ __launch_bounds__(kT) void k_col_pre(const ColStats *__restrict__ cols, int64_t m,
                                                const double *__restrict__ run,
                                                double *__restrict__ pre) {
  __shared__ double wt[kT / 64][kColFields];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j0 = (int64_t)blockIdx.x * kColRun + (int64_t)threadIdx.x * kColPer;
  // this thread's total, then an exclusive scan of the totals over the block
  double tot[kColFields], acc[kColFields];
#pragma unroll
  for (int k = 0; k < kColFields; ++k) tot[k] = 0.0;
#pragma unroll
  for (int i = 0; i < kColPer; ++i) {
    if (j0 + i < m) {
      double f[kColFields];
      col_fields(cols[j0 + i], f);
#pragma unroll
      for (int k = 0; k < kColFields; ++k) tot[k] += f[k];
    }
  }
#pragma unroll
  for (int k = 0; k < kColFields; ++k) {
    double x = tot[k];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double y = __shfl_up(x, o, kWave);
      if (lane >= o) x += y;
    }
    acc[k] = x - tot[k];
    if (lane == 63) wt[w][k] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kColFields; ++k) {
    double before = run[(int64_t)blockIdx.x * kColFields + k];
    for (int i = 0; i < w; ++i) before += wt[i][k];
    acc[k] += before;
  }
#pragma unroll
  for (int i = 0; i < kColPer; ++i) {
    if (j0 + i < m) {
      double f[kColFields];
      col_fields(cols[j0 + i], f);
#pragma unroll
      for (int k = 0; k < kColFields; ++k) {
        pre[(j0 + i) * kColFields + k] = acc[k];
        acc[k] += f[k];
      }
    }
  }
}

// ---- features ----------------------------------------------------------------
// One wave per item.  Rows are the item's lanes; the q range is described by
// the column statistics of the position blocks it meets (their sums: two
// lookups in the prefix sums).  Extents enter as log(extent)/10, centres as
// the position inside the scene's quantisation range (0..1).
// With scene tables (batched scenes) the level range and axes are those of the
// item's scene, and the item count is read from the device.
//
// The wave reductions of an item leave kStat uniform statistics; they are
// parked in the wave's LDS and, once kPark items are parked (or the wave's
// run ends), lane j turns item j's statistics into its 32 features: the
// transcendental part (logs, exps, roots, divisions: several hundred
// instructions) runs once per kPark items instead of once per item.
enum StatField {
  kSNr, kSCandLo, kSCandHi, kSSpan, kSMle1, kSVle1, kSMle2, kSVle2, kSMc1, kSVc1, kSMc2, kSVc2,
  kSFv, kSFe, kSFf, kSMchain, kSN1Lo, kSN1Hi, kSN2Lo, kSN2Hi,
  // column sums over the item's position blocks (ColStats order)
  kSCol, kStat = kSCol + 12
};
// LDS words per parked item: an odd pitch, so the lanes reading one field of
// their items hit distinct banks
constexpr int kStatPitch = kStat | 1;
// Items parked per wave before the feature step.
constexpr int kPark = 64;

__device__ __forceinline__ double stat_f64(const float *s, int lo) {
  const uint64_t u = (uint64_t)(uint32_t)__float_as_int(s[lo]) |
                     ((uint64_t)(uint32_t)__float_as_int(s[lo + 1]) << 32);
  return __builtin_bit_cast(double, u);
}

// Features of one item from its parked statistics.
__device__ __forceinline__ void item_features(const float *__restrict__ st, float *__restrict__ f,
                                              uint32_t *__restrict__ cand_out) {
  const float nr = st[kSNr];
  const long long cand_it = (long long)((uint64_t)(uint32_t)__float_as_int(st[kSCandLo]) |
                                        ((uint64_t)(uint32_t)__float_as_int(st[kSCandHi]) << 32));
  const int span = __float_as_int(st[kSSpan]);
  const float m_le1 = st[kSMle1], v_le1 = st[kSVle1], m_le2 = st[kSMle2], v_le2 = st[kSVle2];
  const float m_c1 = st[kSMc1], v_c1 = st[kSVc1], m_c2 = st[kSMc2], v_c2 = st[kSVc2];
  const float f_v = st[kSFv], f_e = st[kSFe], f_f = st[kSFf], m_chain = st[kSMchain];
  const float *cs = st + kSCol;  // le1, le1sq, le2, le2sq, c1, c1sq, c2, c2sq, nv, ne, nf, cnt
  const float ic = 1.f / fmaxf(cs[11], 1.f);
  const float m_se1 = cs[0] * ic, v_se1 = cs[1] * ic - m_se1 * m_se1;
  const float m_se2 = cs[2] * ic, v_se2 = cs[3] * ic - m_se2 * m_se2;
  const float m_sc1 = cs[4] * ic, v_sc1 = cs[5] * ic - m_sc1 * m_sc1;
  const float m_sc2 = cs[6] * ic, v_sc2 = cs[7] * ic - m_sc2 * m_sc2;
  const float g_v = cs[8] * ic, g_e = cs[9] * ic, g_f = cs[10] * ic;
  const double n1 = stat_f64(st, kSN1Lo), n2 = stat_f64(st, kSN2Lo);
  const float steps = 64.f * (float)(span + 1);
  const float e1r = expf(10.f * m_le1) * n1, e1q = expf(10.f * m_se1) * n1;
  const float e2r = expf(10.f * m_le2) * n2, e2q = expf(10.f * m_se2) * n2;
  const float sd_c1 = sqrtf(fmaxf(v_c1, 0.f)) + sqrtf(fmaxf(v_sc1, 0.f));
  const float sd_c2 = sqrtf(fmaxf(v_c2, 0.f)) + sqrtf(fmaxf(v_sc2, 0.f));
  const float py = fminf(1.f, (e1r + e1q) / (sd_c1 + fabsf(m_c1 - m_sc1) + 1e-12f));
  const float pz = fminf(1.f, (e2r + e2q) / (sd_c2 + fabsf(m_c2 - m_sc2) + 1e-12f));
  const float tc = f_v * g_f + f_f * g_v + f_e * g_e;
  float4 *f4 = reinterpret_cast<float4 *>(f);
  f4[0] = make_float4(nr / 64.f, log1pf((float)cand_it) * 0.1f, log1pf(steps) * 0.1f,
                      (float)cand_it / steps);
  f4[1] = make_float4(m_le1, sqrtf(fmaxf(v_le1, 0.f)), m_le2, sqrtf(fmaxf(v_le2, 0.f)));
  f4[2] = make_float4(m_c1, sqrtf(fmaxf(v_c1, 0.f)), m_c2, sqrtf(fmaxf(v_c2, 0.f)));
  f4[3] = make_float4(f_v, f_e, f_f, m_se1);
  f4[4] = make_float4(sqrtf(fmaxf(v_se1, 0.f)), m_se2, sqrtf(fmaxf(v_se2, 0.f)), m_sc1);
  f4[5] = make_float4(sqrtf(fmaxf(v_sc1, 0.f)), m_sc2, sqrtf(fmaxf(v_sc2, 0.f)), g_v);
  f4[6] = make_float4(g_e, g_f, logf(py + 1e-9f) * 0.1f, logf(pz + 1e-9f) * 0.1f);
  f4[7] = make_float4(tc, log1pf((float)cand_it * py * pz * tc) * 0.1f, log1pf(m_chain) * 0.1f,
                      1.f);
  *cand_out = (uint32_t)min(cand_it, (long long)UINT_MAX);
}
__global__ __launch_bounds__(kT) void k_features(const Item *__restrict__ items, int64_t nitems,
                                                 const int64_t *__restrict__ nitems_dev,
                                                 const int2 *__restrict__ rowlist,
                                                 const uint2 *__restrict__ rowfeat,
                                                 const double *__restrict__ pre, QParams qp0,
                                                 int a10, int a20,
                                                 const SceneParams *__restrict__ sp,
                                                 const int64_t *__restrict__ seg_off, int nscenes,
                                                 float *__restrict__ feat,
                                                 uint32_t *__restrict__ cand) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kT / 64) + (threadIdx.x >> 6)));
  const int nwaves = (int)gridDim.x * (kT / 64);
  if (nitems_dev) nitems = *nitems_dev;
  // Each wave takes a contiguous run of items.  A group's items (its body
  // chunks, then its tail chunks) are consecutive and share their rows, so
  // the row statistics are computed once per run of items with equal rows.
  const int64_t per = (nitems + nwaves - 1) / nwaves;
  const int64_t ib = (int64_t)wave * per;
  const int64_t ie = ib + per < nitems ? ib + per : nitems;
  __shared__ float park_mem[kT / 64][kPark * kStatPitch];
  float *park = park_mem[threadIdx.x >> 6];
  int nslot = 0;       // items parked
  int64_t pbase = ib;  // item of slot 0
  auto flush = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_wave_barrier();
    if (lane < nslot)
      item_features(park + lane * kStatPitch, feat + (pbase + lane) * kFeat, cand + pbase + lane);
    __builtin_amdgcn_wave_barrier();
    pbase += nslot;
    nslot = 0;
  };
  int prow = -1, pnrows = -1;  // rows of the cached statistics
  int64_t sc_lo = 0, sc_hi = -1;  // sorted positions of the cached scene
  bool act = false;
  int p = 0, h = 0;            // this lane's row and its chain end
  double n1 = 0, n2 = 0;
  float nr = 1.f, m_le1 = 0, m_le2 = 0, v_le1 = 0, v_le2 = 0, m_c1 = 0, m_c2 = 0, v_c1 = 0,
        v_c2 = 0, f_v = 0, f_e = 0, f_f = 0, m_chain = 0;
  // The next item and this item's prefix-sum lookups are loaded ahead of
  // the dependent row loads, so their latencies overlap.
  Item nxt = ib < ie ? items[ib] : Item{};
  for (int64_t it = ib; it < ie; ++it) {
    const Item item = nxt;
    if (it + 1 < ie) nxt = items[it + 1];
    const int nrows = item.nrows & kRowsMask;  // tail items: their rows
    const int q0 = item.q0, q1 = item.q1;
    // column sums over the blocks [q0 / kColBlock, q1 / kColBlock]: lane k < 12
    // takes field k
    const int64_t j0 = q0 / kColBlock, j1 = (int64_t)(q1 / kColBlock) + 1;
    double phi = 0.0, plo = 0.0;
    if (lane < kColFields && j1 > j0) {
      phi = pre[j1 * kColFields + lane];
      plo = pre[j0 * kColFields + lane];
    }
    if (item.rowoff != prow || nrows != pnrows) {
      prow = item.rowoff;
      pnrows = nrows;
      // The scene's level range.  A group's rows lie in its planning window,
      // and windows never straddle scenes, so the scene is that of sorted
      // position rowoff; a wave's items are consecutive, so the scene is
      // looked up only when the position leaves the last one's range.
      if (sp ? (item.rowoff < sc_lo || item.rowoff >= sc_hi) : sc_hi < 0) {
        QParams qp = qp0;
        int a1 = a10, a2 = a20;
        if (sp) {
          const int sc = seg_off ? scene_of(seg_off, nscenes, item.rowoff) : 0;
          sc_lo = seg_off ? seg_off[sc] : 0;
          sc_hi = seg_off ? seg_off[sc + 1] : INT64_MAX;
          const SceneParams scp = sp[sc];
          qp = scp.qp;
          a1 = scp.axis == 0 ? 1 : 0;
          a2 = scp.axis == 2 ? 1 : 2;
        } else {
          sc_lo = 0;
          sc_hi = INT64_MAX;
        }
        // axis-indexed fields by selects (a runtime index would put qp in scratch)
        auto pick = [](const double *v, int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); };
        n1 = pick(qp.scale, a1) / kLevelMax;
        n2 = pick(qp.scale, a2) / kLevelMax;
      }
      act = lane < nrows;
      float fr = act ? 1.f : 0.f;
      float le1 = 0, le2 = 0, c1 = 0, c2 = 0, tv = 0, te = 0, tf = 0, chain = 0;
      if (act) {  // the rows and their features, in group order (contiguous)
        const int2 row = rowlist[item.rowoff + lane];
        const RowFeat f = row_feat_unpack(rowfeat[item.rowoff + lane]);
        p = row.x;
        h = row.y;
        le1 = f.le1;
        le2 = f.le2;
        c1 = f.c1;
        c2 = f.c2;
        tv = f.kind == 0 ? 1.f : 0.f;
        te = f.kind == 1 ? 1.f : 0.f;
        tf = f.kind == 2 ? 1.f : 0.f;
        chain = (float)(h - p);
      }
      nr = fmaxf(wave_sum_dpp(fr), 1.f);
      m_le1 = wave_sum_dpp(le1) / nr;
      m_le2 = wave_sum_dpp(le2) / nr;
      v_le1 = wave_sum_dpp(le1 * le1) / nr - m_le1 * m_le1;
      v_le2 = wave_sum_dpp(le2 * le2) / nr - m_le2 * m_le2;
      m_c1 = wave_sum_dpp(c1) / nr;
      m_c2 = wave_sum_dpp(c2) / nr;
      v_c1 = wave_sum_dpp(c1 * c1) / nr - m_c1 * m_c1;
      v_c2 = wave_sum_dpp(c2 * c2) / nr - m_c2 * m_c2;
      f_v = wave_sum_dpp(tv) / nr;
      f_e = wave_sum_dpp(te) / nr;
      f_f = wave_sum_dpp(tf) / nr;
      m_chain = wave_sum_dpp(chain) / nr;
    }
    long long cnd = 0;
    if (act) {
      const int lo = max(q0, p + 1), up = min(q1, h);
      cnd = up >= lo ? (long long)(up - lo + 1) : 0;
    }
    const int span = q1 - q0;
    const long long cand_it = wave_sum(cnd);
    if (lane < kColFields) park[nslot * kStatPitch + kSCol + lane] = (float)(phi - plo);
    if (lane == 0) {
      float *w = park + nslot * kStatPitch;
      const uint64_t cu = (uint64_t)cand_it;
      const uint64_t n1u = __builtin_bit_cast(uint64_t, n1), n2u = __builtin_bit_cast(uint64_t, n2);
      w[kSNr] = nr;
      w[kSCandLo] = __int_as_float((int)(uint32_t)cu);
      w[kSCandHi] = __int_as_float((int)(uint32_t)(cu >> 32));
      w[kSSpan] = __int_as_float(span);
      w[kSMle1] = m_le1;
      w[kSVle1] = v_le1;
      w[kSMle2] = m_le2;
      w[kSVle2] = v_le2;
      w[kSMc1] = m_c1;
      w[kSVc1] = v_c1;
      w[kSMc2] = m_c2;
      w[kSVc2] = v_c2;
      w[kSFv] = f_v;
      w[kSFe] = f_e;
      w[kSFf] = f_f;
      w[kSMchain] = m_chain;
      w[kSN1Lo] = __int_as_float((int)(uint32_t)n1u);
      w[kSN1Hi] = __int_as_float((int)(uint32_t)(n1u >> 32));
      w[kSN2Lo] = __int_as_float((int)(uint32_t)n2u);
      w[kSN2Hi] = __int_as_float((int)(uint32_t)(n2u >> 32));
    }
    if (++nslot == kPark) flush();
  }
  if (nslot) flush();
}

// ---- MLP forward -----------------------------------------------------------------

// Inference on the matrix cores: each wave takes 16 items at a time and runs
// both hidden layers as FP32 GEMMs with v_mfma_f32_16x16x4f32
// (D[16x16] += A[16x4] B[4x16]; lane l holds A[l%16][l/16], B[l/16][l%16]
// and D[4(l/16)+r][l%16], r = 0..3):
//   H1[16 x 64] = relu(X[16 x 32] W1^T + b1): 4 column tiles x 8 k-steps,
//   H2[16 x 64] = relu(H1 W2^T + b2):        4 column tiles x 16 k-steps,
// with H1 passed from the D layout to the A layout through the wave's LDS
// tile, then out = H2 . w3 + b3 as a sum over the 16 lanes of each quarter.
// A lane's B operands (its 96 weights), biases and w3 entries are the same for
// every group, so they are loaded into registers once per wave.
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int kMfmaItems = 16;
constexpr int kH1Row = kHid + 2;  // LDS row stride of the H1 tile: a half wave reads 32 banks

__global__ __launch_bounds__(kT) void k_mlp_forward(const float *__restrict__ params,
                                                    const float *__restrict__ feat, int64_t n,
                                                    const int64_t *__restrict__ n_dev,
                                                    float *__restrict__ pred) {
  __shared__ float h1s[kT / 64][kMfmaItems][kH1Row];
  if (n_dev) n = *n_dev;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, quad = lane >> 4;
  const int64_t ngroups = (n + kMfmaItems - 1) / kMfmaItems;
  const int64_t gstride = (int64_t)gridDim.x * (kT / 64);
  int64_t g = (int64_t)blockIdx.x * (kT / 64) + wv;
  if (g >= ngroups) return;
  float w1r[kHid / 16][kFeat / 4], w2r[kHid / 16][kHid / 4], b1r[kHid / 16], b2r[kHid / 16],
      w3r[kHid / 16];
#pragma unroll
  for (int nt = 0; nt < kHid / 16; ++nt) {
    const int j = 16 * nt + col;
#pragma unroll
    for (int ks = 0; ks < kFeat / 4; ++ks) w1r[nt][ks] = params[kOffW1 + j * kFeat + 4 * ks + quad];
#pragma unroll
    for (int ks = 0; ks < kHid / 4; ++ks) w2r[nt][ks] = params[kOffW2 + j * kHid + 4 * ks + quad];
    b1r[nt] = params[kOffB1 + j];
    b2r[nt] = params[kOffB2 + j];
    w3r[nt] = params[kOffW3 + j];
  }
  const float b3 = params[kOffB3];
  float(*tile)[kH1Row] = h1s[wv];
  // A operand of layer 1: item i0 + col, features 4 ks + quad; the next
  // group's features are fetched while this group runs
  float xn[kFeat / 4];
  auto fetch = [&](int64_t gg) __attribute__((always_inline)) {
    const int64_t ia = gg * kMfmaItems + col;
#pragma unroll
    for (int ks = 0; ks < kFeat / 4; ++ks) xn[ks] = ia < n ? feat[ia * kFeat + 4 * ks + quad] : 0.f;
  };
  fetch(g);
  for (; g < ngroups; g += gstride) {
    const int64_t i0 = g * kMfmaItems;
    float xa[kFeat / 4];
#pragma unroll
    for (int ks = 0; ks < kFeat / 4; ++ks) xa[ks] = xn[ks];
    if (g + gstride < ngroups) fetch(g + gstride);
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kFeat / 4; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[ks], w1r[nt][ks], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) tile[4 * quad + r][16 * nt + col] = fmaxf(acc[r] + b1r[nt], 0.f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // A operand of layer 2: H1[col][4 ks + quad]
    float ha[kHid / 4];
#pragma unroll
    for (int ks = 0; ks < kHid / 4; ++ks) ha[ks] = tile[col][4 * ks + quad];
    // the next group writes the tile only after every lane has read it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kHid / 4; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ha[ks], w2r[nt][ks], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) part[r] += w3r[nt] * fmaxf(acc[r] + b2r[nt], 0.f);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = part[r];
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
      const int64_t item = i0 + 4 * quad + r;
      if (col == 0 && item < n) pred[item] = v + b3;
    }
  }
}

// ---- MLP gradient on the matrix cores ---------------------------------------------
// One block of four waves per 64 minibatch rows, one 16-row group per wave;
// every product is an FP32 v_mfma_f32_16x16x4f32 (lane l holds A[l%16][l/16],
// B[l/16][l%16] and D[4(l/16)+r][l%16], r = 0..3, as in the forward):
//   forward   H1 = relu(X W1^T + b1), H2 = relu(H1 W2^T + b2), y = H2 w3 + b3
//   backward  dH2 = dy w3^T * [H2 > 0],  dH1 = (dH2 W2) * [H1 > 0]
//   weights   dW2 = dH2^T H1, dW1 = dH1^T X (the rows are the K dimension)
// H1 and dH2 go from the D layout to the A layout through a small LDS tile.
// The weight products need no transpose: step s of their K loop feeds row
// 4 (l/16) + s of the group from the D-layout register r = s to both
// operands, so each of the 16 rows is summed once.  Rows are gathered from
// the source rows directly: minibatch row k is source row
// ((start + k) * stride) mod n.  The four waves' gradients are summed in LDS
// in a fixed order ((w0 + w2) + (w1 + w3)) and the block writes one partial
// record (kParams gradients, then the loss).
constexpr int kGradWaves = STQ_GRAD_WAVES;  // waves per block, one 16-row group each
constexpr int kGradRows = 16 * kGradWaves;   // minibatch rows per block
constexpr int kGradRec = kParams + 1;    // a partial record: gradients, loss
constexpr int kTileRow = kHid + 2;       // LDS row stride of a transpose tile
// The weights' B operands in LDS in lane order, [nt][ks][lane], so every
// operand read is 64 consecutive floats: W1 as in layer 1, W2 as in layer 2,
// and W2 as the backward product dH2 W2 reads it.
constexpr int kB1 = 0, kB2 = kB1 + kHid * kFeat, kB2t = kB2 + kHid * kHid,
              kBias = kB2t + kHid * kHid,  // b1, b2, w3, b3 (as in the parameter vector)
              kTiles = kBias + (kParams - kOffB2 + kHid),  // b1 .. b3 end
              kGradOps = kTiles + kGradWaves * 16 * kTileRow,
              kGradLds = kGradOps > kGradWaves * kGradRec ? kGradOps : kGradWaves * kGradRec;

__global__ __launch_bounds__(64 * kGradWaves) void k_mlp_grad(const float *__restrict__ params,
                                                 const float *__restrict__ feat,
                                                 const float *__restrict__ target, int64_t n,
                                                 int64_t start, int64_t stride, int64_t bs,
                                                 float *__restrict__ partial) {
  // weight operands and transpose tiles (16 x kTileRow floats per wave) while
  // computing, then two partial records for the reduction over the waves
  __shared__ float lds[kGradLds];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, quad = lane >> 4;
  float(*tile)[kTileRow] = reinterpret_cast<float(*)[kTileRow]>(lds + kTiles + wv * 16 * kTileRow);
  // coalesced reads of W1 and W2, every load of a thread in flight at once,
  // then scattered into the operand order
  {
    constexpr int NT = 64 * kGradWaves;
    constexpr int N1 = kHid * kFeat / NT, N2 = kHid * kHid / NT;
    static_assert(N1 * NT == kHid * kFeat && N2 * NT == kHid * kHid, "whole loads per thread");
    float x1[N1], x2[N2];
#pragma unroll
    for (int u = 0; u < N1; ++u) x1[u] = params[kOffW1 + threadIdx.x + u * NT];
#pragma unroll
    for (int u = 0; u < N2; ++u) x2[u] = params[kOffW2 + threadIdx.x + u * NT];
#pragma unroll
    for (int u = 0; u < N1; ++u) {
      const int i = threadIdx.x + u * NT, j = i / kFeat, k = i % kFeat;
      lds[kB1 + ((j >> 4) * (kFeat / 4) + (k >> 2)) * 64 + (k & 3) * 16 + (j & 15)] = x1[u];
    }
#pragma unroll
    for (int u = 0; u < N2; ++u) {
      const int i = threadIdx.x + u * NT, j = i / kHid, k = i % kHid;
      lds[kB2 + ((j >> 4) * (kHid / 4) + (k >> 2)) * 64 + (k & 3) * 16 + (j & 15)] = x2[u];
      lds[kB2t + ((k >> 4) * (kHid / 4) + (j >> 2)) * 64 + (j & 3) * 16 + (k & 15)] = x2[u];
    }
  }
  // b1, then b2, w3 and b3 (W2 skipped)
  for (int i = threadIdx.x; i < kHid + (kParams - kOffB2); i += 64 * kGradWaves)
    lds[kBias + i] = params[i < kHid ? kOffB1 + i : kOffB2 + (i - kHid)];
  __syncthreads();
  const float *b1s = lds + kBias, *b2s = lds + kBias + kHid, *w3s = lds + kBias + 2 * kHid;
  const float b3 = lds[kBias + 3 * kHid];
  const float invn = 1.f / (float)bs;
  const int64_t k0 = (int64_t)blockIdx.x * kGradRows + wv * 16;
  // ((start + k) * stride) mod n: the product stays below 2^53, so its double
  // is exact and the double quotient is off by at most one
  const double dn = (double)n;
  auto src = [&](int64_t k) {
    const int64_t a = (start + k) * stride;
    int64_t r = a - (int64_t)((double)a / dn) * n;
    r += r < 0 ? n : 0;
    r -= r >= n ? n : 0;
    return r;
  };

  // X in the A layout (row col, features 4 ks + quad) and in the D layout
  // (rows 4 quad + s, features 16 nt + col), with the targets of those rows
  float xa[kFeat / 4], xd[2][4], td[4];
  bool valid[4];
  {
    const int64_t ka = k0 + col;
    const int64_t ia = ka < bs ? src(ka) : 0;
#pragma unroll
    for (int ks = 0; ks < kFeat / 4; ++ks) xa[ks] = ka < bs ? feat[ia * kFeat + 4 * ks + quad] : 0.f;
#pragma unroll
    for (int sr = 0; sr < 4; ++sr) {
      const int64_t kd = k0 + 4 * quad + sr;
      valid[sr] = kd < bs;
      const int64_t id = valid[sr] ? src(kd) : 0;
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) xd[nt][sr] = valid[sr] ? feat[id * kFeat + 16 * nt + col] : 0.f;
      td[sr] = valid[sr] ? target[id] : 0.f;
    }
  }

  // layer 1: H1 (D layout), passed to the A layout through the tile; the K
  // loop is outermost so the four column tiles' MFMAs are independent
  float h1[kHid / 16][4];
  {
    floatx4 acc[kHid / 16] = {};
#pragma unroll
    for (int ks = 0; ks < kFeat / 4; ++ks)
#pragma unroll
      for (int nt = 0; nt < kHid / 16; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
            xa[ks], lds[kB1 + (nt * (kFeat / 4) + ks) * 64 + lane], acc[nt], 0, 0, 0);
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      const int j = 16 * nt + col;
      const float bj = b1s[j];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        h1[nt][r] = fmaxf(acc[nt][r] + bj, 0.f);
        tile[4 * quad + r][j] = h1[nt][r];
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float ha[kHid / 4];
#pragma unroll
  for (int ks = 0; ks < kHid / 4; ++ks) ha[ks] = tile[col][4 * ks + quad];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // layer 2 and the output
  float h2[kHid / 16][4], w3[kHid / 16], part[4] = {0.f, 0.f, 0.f, 0.f};
  {
    floatx4 acc[kHid / 16] = {};
#pragma unroll
    for (int ks = 0; ks < kHid / 4; ++ks)
#pragma unroll
      for (int nt = 0; nt < kHid / 16; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
            ha[ks], lds[kB2 + (nt * (kHid / 4) + ks) * 64 + lane], acc[nt], 0, 0, 0);
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      const int j = 16 * nt + col;
      const float bj = b2s[j];
      w3[nt] = w3s[j];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        h2[nt][r] = fmaxf(acc[nt][r] + bj, 0.f);
        part[r] += w3[nt] * h2[nt][r];
      }
    }
  }
  float dy[4], loss = 0.f;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float y = part[r];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) y += __shfl_xor(y, o, kWave);  // the 16 lanes of the quad
    const float err = valid[r] ? y + b3 - td[r] : 0.f;
    dy[r] = 2.f * err * invn;
    loss += err * err * invn;
  }

  // backward: dH2 (D layout), then through the tile to the A layout
  float dh2[kHid / 16][4];
#pragma unroll
  for (int nt = 0; nt < kHid / 16; ++nt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      dh2[nt][r] = h2[nt][r] > 0.f ? dy[r] * w3[nt] : 0.f;
      tile[4 * quad + r][16 * nt + col] = dh2[nt][r];
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float da[kHid / 4];
#pragma unroll
  for (int ks = 0; ks < kHid / 4; ++ks) da[ks] = tile[col][4 * ks + quad];
  // dH1 = (dH2 W2) * [H1 > 0], with B[k = j][n = kk] = W2[j][kk]
  float dh1[kHid / 16][4];
  {
    floatx4 acc[kHid / 16] = {};
#pragma unroll
    for (int ks = 0; ks < kHid / 4; ++ks)
#pragma unroll
      for (int nt = 0; nt < kHid / 16; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
            da[ks], lds[kB2t + (nt * (kHid / 4) + ks) * 64 + lane], acc[nt], 0, 0, 0);
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt)
#pragma unroll
      for (int r = 0; r < 4; ++r) dh1[nt][r] = h1[nt][r] > 0.f ? acc[nt][r] : 0.f;
  }

  // weight gradients over the group's 16 rows (step s: rows 4 quad + s)
  float gw2[kHid / 16][kHid / 16][4], gw1[kHid / 16][kFeat / 16][4];
  {
    // the step loop outermost: the 24 tiles' MFMAs are independent
    floatx4 a2[kHid / 16][kHid / 16] = {}, a1[kHid / 16][kFeat / 16] = {};
#pragma unroll
    for (int sr = 0; sr < 4; ++sr)
#pragma unroll
      for (int mt = 0; mt < kHid / 16; ++mt) {
#pragma unroll
        for (int nt = 0; nt < kHid / 16; ++nt)
          a2[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(dh2[mt][sr], h1[nt][sr], a2[mt][nt], 0, 0, 0);
#pragma unroll
        for (int nt = 0; nt < kFeat / 16; ++nt)
          a1[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(dh1[mt][sr], xd[nt][sr], a1[mt][nt], 0, 0, 0);
      }
#pragma unroll
    for (int mt = 0; mt < kHid / 16; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int nt = 0; nt < kHid / 16; ++nt) gw2[mt][nt][r] = a2[mt][nt][r];
#pragma unroll
        for (int nt = 0; nt < kFeat / 16; ++nt) gw1[mt][nt][r] = a1[mt][nt][r];
      }
  }
  // biases, w3, b3 and the loss: sums over the rows (r, then the quads)
  float gb1[kHid / 16], gb2[kHid / 16], gw3[kHid / 16];
#pragma unroll
  for (int nt = 0; nt < kHid / 16; ++nt) {
    float s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      s1 += dh1[nt][r];
      s2 += dh2[nt][r];
      s3 += dy[r] * h2[nt][r];
    }
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      s1 += __shfl_xor(s1, o, kWave);
      s2 += __shfl_xor(s2, o, kWave);
      s3 += __shfl_xor(s3, o, kWave);
    }
    gb1[nt] = s1;
    gb2[nt] = s2;
    gw3[nt] = s3;
  }
  float gb3 = dy[0] + dy[1] + dy[2] + dy[3];
#pragma unroll
  for (int o = 16; o < 64; o <<= 1) {
    gb3 += __shfl_xor(gb3, o, kWave);
    loss += __shfl_xor(loss, o, kWave);
  }

  // every entry of the record has one owning lane; the same visiting order
  // serves the LDS exchange and the global store
  auto visit = [&](auto &&f) __attribute__((always_inline)) {
#pragma unroll
    for (int mt = 0; mt < kHid / 16; ++mt) {
#pragma unroll
      for (int nt = 0; nt < kHid / 16; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          f(kOffW2 + (16 * mt + 4 * quad + r) * kHid + 16 * nt + col, gw2[mt][nt][r]);
#pragma unroll
      for (int nt = 0; nt < kFeat / 16; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          f(kOffW1 + (16 * mt + 4 * quad + r) * kFeat + 16 * nt + col, gw1[mt][nt][r]);
    }
    if (quad == 0) {
#pragma unroll
      for (int nt = 0; nt < kHid / 16; ++nt) {
        f(kOffB1 + 16 * nt + col, gb1[nt]);
        f(kOffB2 + 16 * nt + col, gb2[nt]);
        f(kOffW3 + 16 * nt + col, gw3[nt]);
      }
      if (col == 0) {
        f(kOffB3, gb3);
        f(kParams, loss);
      }
    }
  };
  // each wave's record into its own LDS record, then the block sums them in a
  // fixed tree order (four waves: (w0 + w2) + (w1 + w3)) and writes the
  // partial record with consecutive threads on consecutive entries
  __syncthreads();  // operands and tiles are done with: the records reuse the LDS
  visit([&](int i, float &v) { lds[wv * kGradRec + i] = v; });
  __syncthreads();
  float *out = partial + (int64_t)blockIdx.x * kGradRec;
  for (int i = threadIdx.x; i < kGradRec; i += 64 * kGradWaves) {
    float v[kGradWaves];
#pragma unroll
    for (int w = 0; w < kGradWaves; ++w) v[w] = lds[w * kGradRec + i];
#pragma unroll
    for (int h = kGradWaves / 2; h >= 1; h /= 2)
#pragma unroll
      for (int w = 0; w < h; ++w) v[w] += v[w + h];
    out[i] = v[0];
  }
}

// Sum of the partial records of nb blocks in a fixed order: thread (p, g) of
// a block sums records g, g + 16, ... of entry p (64 consecutive entries per
// wave, so each record is read in whole lines; eight loads in flight per
// thread), then the 16 sums are added in order.  With ADAM the sum is applied as one Adam step to the
// parameters; otherwise it is stored as the gradient (and the loss).
constexpr int kRedP = 64, kRedG = 16;
template <bool ADAM>
__global__ __launch_bounds__(kRedP * kRedG) void k_grad_reduce(const float *__restrict__ partial, int nb,
                                                    float *__restrict__ grad,
                                                    float *__restrict__ loss,
                                                    float *__restrict__ w, float *__restrict__ m,
                                                    float *__restrict__ v, float lr, float c1,
                                                    float c2) {
  __shared__ float sums[kRedG][kRedP];
  const int pl = threadIdx.x % kRedP, g = threadIdx.x / kRedP;
  const int p = blockIdx.x * kRedP + pl;
  float s = 0.f;
  if (p < kGradRec) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int b = g;
    for (; b + 7 * kRedG < nb; b += 8 * kRedG) {
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += partial[(int64_t)(b + u * kRedG) * kGradRec + p];
    }
    for (int u = 0; b < nb; b += kRedG, ++u) acc[u & 7] += partial[(int64_t)b * kGradRec + p];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += acc[u];
  }
  sums[g][pl] = s;
  __syncthreads();
  if (g != 0 || p >= kGradRec) return;
  float gi = sums[0][pl];
  for (int k = 1; k < kRedG; ++k) gi += sums[k][pl];
  if (!ADAM) {
    if (p < kParams) grad[p] = gi;
    else *loss = gi;
    return;
  }
  if (p >= kParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mi = b1 * m[p] + (1.f - b1) * gi;
  const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  w[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(kT) void k_caps(const float *__restrict__ pred,
                                             const uint32_t *__restrict__ cand, int64_t n,
                                             const int64_t *__restrict__ n_dev,
                                             float alpha, float beta,
                                             uint32_t *__restrict__ caps) {
  if (n_dev) n = *n_dev;  // items beyond the device-side count are never read
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const float pr = fminf(fmaxf(pred[i], 0.f), 30.f);
    const float want = ceilf(alpha * expm1f(pr) + beta);
    const uint32_t c = cand[i];
    caps[i] = want >= (float)c ? c : (uint32_t)want;
  }
}

__global__ __launch_bounds__(kT) void k_cap_budget(uint32_t *__restrict__ caps,
                                                   const int64_t *__restrict__ cap_off, int64_t n,
                                                   const int64_t *__restrict__ n_dev,
                                                   int64_t budget) {
  if (n_dev) n = *n_dev;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    if (cap_off[i] + (int64_t)caps[i] > budget) caps[i] = 0;
}

__global__ __launch_bounds__(kT) void k_targets(const uint32_t *__restrict__ counts, int64_t n,
                                                float *__restrict__ target) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) target[i] = log1pf((float)counts[i]);
}

__global__ __launch_bounds__(kT) void k_abs_error(const float *__restrict__ pred,
                                                  const uint32_t *__restrict__ counts, int64_t n,
                                                  const int64_t *__restrict__ n_dev,
                                                  double *__restrict__ acc) {
  if (n_dev) n = *n_dev;
  double e = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    e += fabs((double)expm1f(fminf(fmaxf(pred[i], 0.f), 30.f)) - (double)counts[i]);
  block_atomic_add<kT>(e, acc);
}

__global__ __launch_bounds__(kT) void k_compact(const uint32_t *__restrict__ counts,
                                                const uint32_t *__restrict__ caps,
                                                const int64_t *__restrict__ cap_off,
                                                const int64_t *__restrict__ off, int64_t n,
                                                const int64_t *__restrict__ n_dev,
                                                const int64_t *__restrict__ total, int64_t capacity,
                                                const int2 *__restrict__ scratch,
                                                int2 *__restrict__ out, int *__restrict__ ovf,
                                                unsigned long long *__restrict__ novf) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kT / 64) + (threadIdx.x >> 6)));
  const int nwaves = (int)gridDim.x * (kT / 64);
  if (total && *total > capacity) return;  // the output does not fit: nothing is written
  if (n_dev) n = *n_dev;
  for (int64_t it = wave; it < n; it += nwaves) {
    const uint32_t c = counts[it];
    if (c <= caps[it]) {
      const int2 *src = scratch + cap_off[it];
      int2 *dst = out + off[it];
      // rounds of up to eight loads in flight per lane (512 pairs), the last
      // round's loads predicated: one memory latency per 512 pairs
      for (uint32_t k = lane; k < c; k += 8 * kWave) {
        int2 x[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (k + r * kWave < c) x[r] = src[k + r * kWave];
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (k + r * kWave < c) dst[k + r * kWave] = x[r];
      }
    } else if (lane == 0) {
      ovf[atomicAdd(novf, 1ull)] = (int)it;
    }
  }
}

__global__ __launch_bounds__(kT) void k_gather_batch(const float *__restrict__ feat,
                                                     const float *__restrict__ target, int64_t n,
                                                     int64_t start, int64_t stride, int64_t bs,
                                                     float *__restrict__ bfeat,
                                                     float *__restrict__ btarget) {
  const int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (k >= bs) return;
  const int64_t i = (int64_t)(((unsigned __int128)(start + k) * (unsigned __int128)stride) % (unsigned __int128)n);
  for (int f = 0; f < kFeat; ++f) bfeat[k * kFeat + f] = feat[i * kFeat + f];
  btarget[k] = target[i];
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

size_t col_work_doubles(int64_t ncols) {
  return (size_t)kColFields * ((size_t)ncols + 1 + (size_t)((ncols + kColRun - 1) / kColRun));
}

void launch_features(bool f32, const SweepArgs &sa, const ColStats *cols, int64_t ncols,
                     double *col_work, const QParams &qp, int axis, const SceneParams *sp,
                     const int64_t *seg_off, int nscenes, float *feat, uint32_t *cand,
                     hipStream_t st) {
  if (sa.nitems <= 0 || ncols <= 0) return;
  // prefix sums of the column statistics: pre[0 .. ncols], then the run sums
  double *pre = col_work;
  double *run = col_work + (size_t)kColFields * ((size_t)ncols + 1);
  const int64_t nr = (ncols + kColRun - 1) / kColRun;
  const unsigned gr = (unsigned)nr;
  hipLaunchKernelGGL(k_col_runs, dim3(gr), dim3(kT), 0, st, cols, ncols, run);
  hipLaunchKernelGGL(k_col_top, dim3(1), dim3(kTopThreads), 0, st, run, nr,
                     pre + (size_t)kColFields * (size_t)ncols);
  hipLaunchKernelGGL(k_col_pre, dim3(gr), dim3(kT), 0, st, cols, ncols, (const double *)run, pre);
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const unsigned g = (unsigned)std::min<int64_t>((sa.nitems + 3) / 4, 4096);
  hipLaunchKernelGGL(k_features, dim3(g), dim3(kT), 0, st, sa.items, sa.nitems, sa.nitems_dev,
                     sa.rowlist, sa.rowfeat, (const double *)pre, qp, a1, a2, sp, seg_off, nscenes,
                     feat, cand);
  STQ_HIP(hipGetLastError());
}

void launch_mlp_forward(const float *params, const float *feat, int64_t n, float *pred,
                        hipStream_t st, const int64_t *n_dev) {
  if (n <= 0) return;
  // 16 items per wave step; one block per CU, each wave looping over groups,
  // so a wave's weight loads (96 per lane) serve many groups
  const int64_t per_block = (kT / 64) * kMfmaItems;
  const unsigned g = (unsigned)std::min<int64_t>((n + per_block - 1) / per_block, 256);
  hipLaunchKernelGGL(k_mlp_forward, dim3(g), dim3(kT), 0, st, params, feat, n, n_dev, pred);
  STQ_HIP(hipGetLastError());
}

void launch_caps(const float *pred, const uint32_t *cand, int64_t n, float alpha, float beta,
                 uint32_t *caps, hipStream_t st, const int64_t *n_dev) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_caps, dim3(std::min(nblk(n), 8192u)), dim3(kT), 0, st, pred, cand, n, n_dev,
                     alpha, beta, caps);
  STQ_HIP(hipGetLastError());
}

void launch_cap_budget(uint32_t *caps, const int64_t *cap_off, int64_t n, int64_t budget,
                       hipStream_t st, const int64_t *n_dev) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cap_budget, dim3(std::min(nblk(n), 8192u)), dim3(kT), 0, st, caps, cap_off, n,
                     n_dev, budget);
  STQ_HIP(hipGetLastError());
}

void launch_targets(const uint32_t *counts, int64_t n, float *targets, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_targets, dim3(nblk(n)), dim3(kT), 0, st, counts, n, targets);
  STQ_HIP(hipGetLastError());
}

size_t mlp_partial_floats(int64_t bs) {
  return (size_t)((bs + kGradRows - 1) / kGradRows) * kGradRec;
}

void launch_mlp_gradient(const float *params, const float *feat, const float *targets, int64_t n,
                         float *partial, float *grad, float *loss, hipStream_t st) {
  if (n <= 0) return;
  const int nb = (int)((n + kGradRows - 1) / kGradRows);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(64 * kGradWaves), 0, st, params, feat, targets, n,
                     (int64_t)0,
                     (int64_t)1, n, partial);
  hipLaunchKernelGGL(k_grad_reduce<false>, dim3((kGradRec + kRedP - 1) / kRedP), dim3(kRedP * kRedG), 0, st,
                     (const float *)partial, nb, grad, loss, nullptr, nullptr, nullptr, 0.f, 1.f, 1.f);
  STQ_HIP(hipGetLastError());
}

void launch_train_step(float *params, float *m, float *v, const float *feat, const float *targets,
                       int64_t n, int64_t start, int64_t stride, int64_t bs, float *partial, float lr,
                       int t, hipStream_t st) {
  if (n <= 0 || bs <= 0) return;
  const int nb = (int)((bs + kGradRows - 1) / kGradRows);
  const float c1 = 1.f - powf(0.9f, (float)t), c2 = 1.f - powf(0.999f, (float)t);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(64 * kGradWaves), 0, st, (const float *)params, feat,
                     targets, n,
                     start % n, stride, bs, partial);
  hipLaunchKernelGGL(k_grad_reduce<true>, dim3((kGradRec + kRedP - 1) / kRedP), dim3(kRedP * kRedG), 0, st,
                     (const float *)partial, nb, nullptr, nullptr, params, m, v, lr, c1, c2);
  STQ_HIP(hipGetLastError());
}

void launch_abs_error(const float *pred, const uint32_t *counts, int64_t n, double *acc,
                      hipStream_t st, const int64_t *n_dev) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_abs_error, dim3(std::min(nblk(n), kReduceGrid)), dim3(kT), 0, st, pred,
                     counts, n, n_dev, acc);
  STQ_HIP(hipGetLastError());
}

void launch_compact(const uint32_t *counts, const uint32_t *caps, const int64_t *cap_off,
                    const int64_t *off, int64_t n, const int2 *scratch, int2 *out, int *ovf,
                    unsigned long long *novf, hipStream_t st, const int64_t *n_dev,
                    const int64_t *total, int64_t capacity) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 3) / 4, 8192);
  hipLaunchKernelGGL(k_compact, dim3(g), dim3(kT), 0, st, counts, caps, cap_off, off, n, n_dev,
                     total, capacity, scratch, out, ovf, novf);
  STQ_HIP(hipGetLastError());
}

void launch_gather_batch(const float *feat, const float *target, int64_t n, int64_t start,
                         int64_t stride, int64_t bs, float *bfeat, float *btarget, hipStream_t st) {
  if (bs <= 0) return;
  hipLaunchKernelGGL(k_gather_batch, dim3(nblk(bs)), dim3(kT), 0, st, feat, target, n, start,
                     stride, bs, bfeat, btarget);
  STQ_HIP(hipGetLastError());
}

// splitmix64 of (seed, index) -> uniform in [-1, 1)
static inline float hash_uniform(uint64_t seed, uint64_t i) {
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (float)((double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
}

void sizer_init_params(float *w, uint64_t seed) {
  const float a1 = std::sqrt(6.f / kFeat), a2 = std::sqrt(6.f / kHid), a3 = std::sqrt(1.f / kHid);
  for (int i = 0; i < kParams; ++i) w[i] = 0.f;
  for (int i = 0; i < kHid * kFeat; ++i) w[kOffW1 + i] = a1 * hash_uniform(seed, kOffW1 + i);
  for (int i = 0; i < kHid * kHid; ++i) w[kOffW2 + i] = a2 * hash_uniform(seed, kOffW2 + i);
  for (int i = 0; i < kHid; ++i) w[kOffW3 + i] = a3 * hash_uniform(seed, kOffW3 + i);
  w[kOffB3] = 2.f;
}

}  // namespace stqamd
