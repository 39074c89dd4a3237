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
// Per axis: sum(min+max), sum(min^2+max^2), sum(max-min), sum(((min+max)/2)^2).
__global__ __launch_bounds__(kThreads) void k_axis_stats(const stq_aabb_f64 *__restrict__ boxes,
                                                         int64_t n, double *__restrict__ partial) {
  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const stq_aabb_f64 &b = boxes[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double lo = b.min[a], up = b.max[a];
      const double c = 0.5 * (lo + up);
      acc[4 * a + 0] += lo + up;
      acc[4 * a + 1] += lo * lo + up * up;
      acc[4 * a + 2] += up - lo;
      acc[4 * a + 3] += c * c;
    }
  }
  __shared__ double red[kThreads];
  for (int k = 0; k < 12; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * 12 + k] = red[0];
    __syncthreads();
  }
}

// ---- keys ---------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_keys_f64(const stq_aabb_f64 *__restrict__ boxes,
                                                       int64_t n, int axis,
                                                       uint64_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals,
                                                       unsigned *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
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
                                                       uint32_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals,
                                                       unsigned *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
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

__global__ __launch_bounds__(kThreads) void k_gather_f64(const stq_aabb_f64 *__restrict__ boxes,
                                                         const uint32_t *__restrict__ order,
                                                         int64_t n, int axis,
                                                         double *__restrict__ smin,
                                                         double *__restrict__ smax,
                                                         Rec64 *__restrict__ rec) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (s >= n) return;
  const stq_aabb_f64 &b = boxes[order[s]];
  int a1, a2;
  other_axes(axis, a1, a2);
  smin[s] = b.min[axis];
  smax[s] = b.max[axis];
  Rec64 r;
  r.lo1 = b.min[a1];
  r.hi1 = b.max[a1];
  r.lo2 = b.min[a2];
  r.hi2 = b.max[a2];
  r.v0 = b.vertex_ids[0];
  r.v1 = b.vertex_ids[1];
  r.v2 = b.vertex_ids[2];
  r.type = type_bit(r.v1, r.v2);
  rec[s] = r;
}

__global__ __launch_bounds__(kThreads) void k_gather_f32(const stq_aabb_f64 *__restrict__ boxes,
                                                         const uint32_t *__restrict__ order,
                                                         int64_t n, int axis,
                                                         float *__restrict__ smin,
                                                         float *__restrict__ smax,
                                                         Rec32 *__restrict__ rec,
                                                         Full64 *__restrict__ full) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (s >= n) return;
  const stq_aabb_f64 &b = boxes[order[s]];
  int a1, a2;
  other_axes(axis, a1, a2);
  smin[s] = round_down_f32(b.min[axis]);
  smax[s] = round_up_f32(b.max[axis]);
  Rec32 r;
  r.lo1 = round_down_f32(b.min[a1]);
  r.hi1 = round_up_f32(b.max[a1]);
  r.lo2 = round_down_f32(b.min[a2]);
  r.hi2 = round_up_f32(b.max[a2]);
  r.v0 = b.vertex_ids[0];
  r.v1 = b.vertex_ids[1];
  r.v2 = b.vertex_ids[2];
  r.type = type_bit(r.v1, r.v2);
  rec[s] = r;
  Full64 f;
  f.lo0 = b.min[axis];
  f.hi0 = b.max[axis];
  f.lo1 = b.min[a1];
  f.hi1 = b.max[a1];
  f.lo2 = b.min[a2];
  f.hi2 = b.max[a2];
  full[s] = f;
}

// ---- chain ends -----------------------------------------------------------------
// hi[s] = last q with smin[q] <= smax[s] (q >= s since smin[s] <= smax[s]);
// galloping search forward from s, then bisection.  Adds sum(hi - s) to *cand.
template <class T>
__global__ __launch_bounds__(kThreads) void k_chain_end(const T *__restrict__ smin,
                                                        const T *__restrict__ smax, int64_t n,
                                                        int *__restrict__ hi,
                                                        unsigned long long *__restrict__ cand) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  unsigned long long len = 0;
  if (s < n) {
    const T x = smax[s];
    int64_t lo = s, step = 1, up;
    while (true) {
      const int64_t probe = lo + step;
      if (probe >= n) { up = n; break; }
      if (smin[probe] <= x) { lo = probe; step <<= 1; }
      else { up = probe; break; }
    }
    // smin[lo] <= x, and up == n or smin[up] > x
    while (up - lo > 1) {
      const int64_t mid = lo + ((up - lo) >> 1);
      if (smin[mid] <= x) lo = mid; else up = mid;
    }
    hi[s] = (int)lo;
    len = (unsigned long long)(lo - s);
  }
  // wave reduction then one atomic per wave
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) len += __shfl_xor(len, o, kWave);
  if ((threadIdx.x & 63) == 0 && len) atomicAdd(cand, len);
}

// ---- work planning ----------------------------------------------------------------
// One block per window of W sorted rows (W a power of two, 64..2048): rows are
// ordered by (hi, position) with a bitonic sort in LDS and cut into groups of
// 64.  Group g covers rowlist[64g .. 64g+nrows) and the q range
// [min p + 1, max hi].  nchunks[g] = number of chunk-sized items.
constexpr int kMaxWindow = 2048;

__global__ __launch_bounds__(kThreads) void k_plan(const int *__restrict__ hi, int64_t n, int W,
                                                   int chunk, int *__restrict__ rowlist,
                                                   int4 *__restrict__ groups,
                                                   uint32_t *__restrict__ nchunks,
                                                   unsigned long long *__restrict__ lane_steps) {
  __shared__ uint64_t key[kMaxWindow];
  const int64_t wbase = (int64_t)blockIdx.x * W;
  const int cnt = (int)std::min<int64_t>(W, n - wbase);
  for (int j = threadIdx.x; j < W; j += kThreads) {
    uint64_t k = ~0ull;
    if (j < cnt) k = ((uint64_t)(uint32_t)hi[wbase + j] << 32) | (uint32_t)j;
    key[j] = k;
  }
  __syncthreads();
  // bitonic sort ascending of W keys
  for (int size = 2; size <= W; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int j = threadIdx.x; j < W; j += kThreads) {
        const int pj = j ^ stride;
        if (pj > j) {
          const bool up = (j & size) == 0;
          const uint64_t a = key[j], b = key[pj];
          if ((a > b) == up) {
            key[j] = b;
            key[pj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int j = threadIdx.x; j < cnt; j += kThreads)
    rowlist[wbase + j] = (int)(wbase + (int64_t)(key[j] & 0xffffffffu));
  // one wave per group
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ng = (cnt + 63) / 64;
  unsigned long long steps = 0;
  for (int k = w; k < ng; k += kThreads / 64) {
    const int j = 64 * k + lane;
    const bool act = j < cnt;
    int p = act ? (int)(wbase + (int64_t)(key[j] & 0xffffffffu)) : INT_MAX;
    int h = act ? (int)(key[j] >> 32) : INT_MIN;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      p = min(p, __shfl_xor(p, o, kWave));
      h = max(h, __shfl_xor(h, o, kWave));
    }
    const int nrows = min(64, cnt - 64 * k);
    const int q0 = p + 1, q1 = h;
    const int64_t g = (wbase >> 6) + k;
    if (lane == 0) {
      groups[g] = make_int4((int)(wbase + 64 * k), nrows, q0, q1);
      const int64_t span = (int64_t)q1 - q0 + 1;
      nchunks[g] = span > 0 ? (uint32_t)((span + chunk - 1) / chunk) : 0u;
      if (span > 0) steps += (unsigned long long)span * 64ull;
    }
  }
  if (lane == 0 && steps) atomicAdd(lane_steps, steps);
}

__global__ __launch_bounds__(kThreads) void k_items(const int4 *__restrict__ groups,
                                                    const int64_t *__restrict__ ioff, int64_t ngroups,
                                                    int chunk, Item *__restrict__ items) {
  const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (g >= ngroups) return;
  const int4 gr = groups[g];
  const int64_t o = ioff[g], e = ioff[g + 1];
  for (int64_t c = 0; c < e - o; ++c) {
    Item it;
    it.rowoff = gr.x;
    it.nrows = gr.y;
    it.q0 = gr.z + (int)(c * chunk);
    it.q1 = min(it.q0 + chunk - 1, gr.w);
    items[o + c] = it;
  }
}

// ---- the sweep ----------------------------------------------------------------------
// Every test is a v_cmp whose lane mask is combined with s_and: the ballot of
// each comparison is taken separately, so no short-circuit branches appear.
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

template <class R>
__device__ __forceinline__ uint64_t geo_mask(const R &m, const R &q) {
  return ballot(m.hi1 >= q.lo1) & ballot(m.lo1 <= q.hi1) & ballot(m.hi2 >= q.lo2) &
         ballot(m.lo2 <= q.hi2);
}

__device__ __forceinline__ bool confirm64(const Full64 &m, const Full64 &q) {
  return (m.hi0 >= q.lo0) & (m.lo0 <= q.hi0) & (m.hi1 >= q.lo1) & (m.lo1 <= q.hi1) &
         (m.hi2 >= q.lo2) & (m.lo2 <= q.hi2);
}

// FILL = false: counts[it] = pairs of item it.  FILL = true: pairs of item
// it go to out[offsets[it] ..) as (input id of the row, sorted q).
// CONFIRM: R holds FP32 outward bounds; every FP32 hit is re-tested in FP64.
// B: q records per scalar-load batch (the next batch is in flight while the
// current one is tested).
template <class R, bool CONFIRM, bool FILL, int B>
__global__ __launch_bounds__(kThreads) void k_sweep(
    const Item *__restrict__ items, int64_t nitems, const int *__restrict__ rowlist,
    const int *__restrict__ hi, const R *__restrict__ rec, const Full64 *__restrict__ full,
    const uint32_t *__restrict__ orig, uint32_t *__restrict__ counts,
    const int64_t *__restrict__ offsets, int2 *__restrict__ out,
    unsigned long long *__restrict__ confirms, int zero) {
  const int lane = threadIdx.x & 63;
  const int wave = uniform((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  const int nwaves = (int)gridDim.x * (kThreads / 64);
  unsigned long long nconf = 0;
  for (int64_t it = wave; it < nitems; it += nwaves) {
    const Item item = items[it];
    const bool act = lane < item.nrows;
    const int p = act ? rowlist[item.rowoff + lane] : INT_MAX;
    const int myhi = act ? hi[p] : INT_MIN;
    R mine;
    mine.lo1 = mine.lo2 = (decltype(mine.lo1))INFINITY;  // empty box for idle lanes
    mine.hi1 = mine.hi2 = -(decltype(mine.lo1))INFINITY;
    mine.v0 = mine.v1 = mine.v2 = mine.type = 0;
    Full64 mf{};
    uint32_t op = 0;
    if (act) {
      mine = rec[p];
      if (CONFIRM) mf = full[p];
      if (FILL) op = orig[p];
    }
    const int acc = act ? accept_bits(mine.type) : 0;
    unsigned cnt = 0;
    int64_t base = FILL ? offsets[it] : 0;

    // rows in m passed the geometric test against sorted q (record rq)
    auto hit = [&](int q, const R &rq, uint64_t m) __attribute__((always_inline)) {
      const bool h0 = (((m >> lane) & 1ull) != 0) & ((acc & rq.type) != 0) &
                      (mine.v0 != rq.v0) & (mine.v0 != rq.v1) & (mine.v0 != rq.v2) &
                      (mine.v1 != rq.v0) & (mine.v1 != rq.v1) & (mine.v1 != rq.v2) &
                      (mine.v2 != rq.v0) & (mine.v2 != rq.v1) & (mine.v2 != rq.v2);
      bool h = h0;
      if (CONFIRM) {
        const uint64_t hm0 = ballot(h0);
        if (hm0 == 0) return;
        nconf += (unsigned long long)__popcll(hm0);
        // a uniform address the compiler cannot prove uniform: a vector load,
        // so its wait does not drain the scalar prefetch
        const Full64 fq = full[q + zero * lane];
        h = h0 & confirm64(mf, fq);
      }
      const uint64_t hm = ballot(h);
      if (hm == 0) return;
      if (FILL) {
        if (h) out[base + mbcnt64(hm)] = make_int2((int)op, q);
        base += __popcll(hm);
      } else {
        cnt += (unsigned)__popcll(hm);
      }
    };

    // Geometric masks of one batch of B records (inactive lanes can pass the
    // geometric test only against unbounded boxes, and never the type test).
    auto batch = [&](const R *b, int q, bool check) __attribute__((always_inline)) {
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

    // q in [qa, qb].  Scalar loads return out of order, so a wait for one
    // batch waits for every load in flight; the next batch's address is made
    // to depend on the current batch (times the runtime zero) so that its
    // loads are issued after that wait, and it is in flight while the
    // current batch is tested.  Two buffers alternate.
    auto walk = [&](int qa, int qb, bool check) __attribute__((always_inline)) {
      int q = qa;
      int nb = (qb - qa + 1) / B;
      if (nb > 0) {
        R b0[B], b1[B];
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
        const R r = rec[q];
        uint64_t m = geo_mask(mine, r);
        if (check) m &= ballot(q > p) & ballot(q <= myhi);
        if (m) hit(q, r, m);
      }
    };

    // every active lane has p < q <= hi_p for q in (pmax, hmin]
    int pmax = act ? p : INT_MIN, hmin = act ? myhi : INT_MAX;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      pmax = max(pmax, __shfl_xor(pmax, o, kWave));
      hmin = min(hmin, __shfl_xor(hmin, o, kWave));
    }
    pmax = uniform(pmax);
    hmin = uniform(hmin);
    const int q0 = item.q0, q1 = item.q1;
    const int b0 = max(q0, pmax + 1), b1 = min(q1, hmin);
    if (b0 <= b1) {
      if (q0 < b0) walk(q0, b0 - 1, true);
      walk(b0, b1, false);
      if (b1 < q1) walk(b1 + 1, q1, true);
    } else {
      walk(q0, q1, true);
    }
    if (!FILL && lane == 0) counts[it] = cnt;
  }
  if (CONFIRM && !FILL && lane == 0 && nconf) atomicAdd(confirms, nconf);
}

// (input id of the row, sorted q) -> canonical (min id, max id).
__global__ __launch_bounds__(kThreads) void k_translate(int2 *__restrict__ pairs, int64_t count,
                                                        const uint32_t *__restrict__ orig) {
  const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (k >= count) return;
  const int2 e = pairs[k];
  const int oq = (int)orig[e.y];
  pairs[k] = make_int2(min(e.x, oq), max(e.x, oq));
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

// ---- host wrappers -------------------------------------------------------------------
void launch_axis_stats(const stq_aabb_f64 *boxes, int64_t n, double *partial, int nblocks,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_axis_stats, dim3(nblocks), dim3(kThreads), 0, st, boxes, n, partial);
  STQ_HIP(hipGetLastError());
}

void launch_keys(bool f32, const stq_aabb_f64 *boxes, int64_t n, int axis, void *keys,
                 uint32_t *vals, unsigned *err, hipStream_t st) {
  if (n <= 0) return;
  if (f32)
    hipLaunchKernelGGL(k_keys_f32, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, n, axis,
                       (uint32_t *)keys, vals, err);
  else
    hipLaunchKernelGGL(k_keys_f64, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, n, axis,
                       (uint64_t *)keys, vals, err);
  STQ_HIP(hipGetLastError());
}

void launch_gather(bool f32, const stq_aabb_f64 *boxes, const uint32_t *order, int64_t n,
                   int axis, void *smin, void *smax, void *rec, Full64 *full, hipStream_t st) {
  if (n <= 0) return;
  if (f32)
    hipLaunchKernelGGL(k_gather_f32, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, order,
                       n, axis, (float *)smin, (float *)smax, (Rec32 *)rec, full);
  else
    hipLaunchKernelGGL(k_gather_f64, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, order,
                       n, axis, (double *)smin, (double *)smax, (Rec64 *)rec);
  STQ_HIP(hipGetLastError());
}

void launch_chain_end(bool f32, const void *smin, const void *smax, int64_t n, int *hi,
                      unsigned long long *cand, hipStream_t st) {
  if (n <= 0) return;
  if (f32)
    hipLaunchKernelGGL(k_chain_end<float>, dim3(blocks_for(n)), dim3(kThreads), 0, st,
                       (const float *)smin, (const float *)smax, n, hi, cand);
  else
    hipLaunchKernelGGL(k_chain_end<double>, dim3(blocks_for(n)), dim3(kThreads), 0, st,
                       (const double *)smin, (const double *)smax, n, hi, cand);
  STQ_HIP(hipGetLastError());
}

void launch_plan(const int *hi, int64_t n, int W, int chunk, int *rowlist, int4 *groups,
                 uint32_t *nchunks, unsigned long long *lane_steps, hipStream_t st) {
  if (n <= 0) return;
  const int64_t nwin = (n + W - 1) / W;
  hipLaunchKernelGGL(k_plan, dim3((unsigned)nwin), dim3(kThreads), 0, st, hi, n, W, chunk,
                     rowlist, groups, nchunks, lane_steps);
  STQ_HIP(hipGetLastError());
}

void launch_items(const int4 *groups, const int64_t *ioff, int64_t ngroups, int chunk,
                  Item *items, hipStream_t st) {
  if (ngroups <= 0) return;
  hipLaunchKernelGGL(k_items, dim3(blocks_for(ngroups)), dim3(kThreads), 0, st, groups, ioff,
                     ngroups, chunk, items);
  STQ_HIP(hipGetLastError());
}

void launch_sweep(bool f32, bool fill, int nblocks, const Item *items, int64_t nitems,
                  const int *rowlist, const int *hi, const void *rec, const Full64 *full,
                  const uint32_t *orig, uint32_t *counts, const int64_t *offsets, stq_pair *out,
                  unsigned long long *confirms, hipStream_t st) {
  if (nitems <= 0) return;
  const dim3 grid(nblocks), block(kThreads);
  int2 *o = reinterpret_cast<int2 *>(out);
  if (f32) {
    const Rec32 *r = (const Rec32 *)rec;
    if (fill)
      hipLaunchKernelGGL((k_sweep<Rec32, true, true, 3>), grid, block, 0, st, items, nitems,
                         rowlist, hi, r, full, orig, counts, offsets, o, confirms, 0);
    else
      hipLaunchKernelGGL((k_sweep<Rec32, true, false, 3>), grid, block, 0, st, items, nitems,
                         rowlist, hi, r, full, orig, counts, offsets, o, confirms, 0);
  } else {
    const Rec64 *r = (const Rec64 *)rec;
    if (fill)
      hipLaunchKernelGGL((k_sweep<Rec64, false, true, 2>), grid, block, 0, st, items, nitems,
                         rowlist, hi, r, full, orig, counts, offsets, o, confirms, 0);
    else
      hipLaunchKernelGGL((k_sweep<Rec64, false, false, 2>), grid, block, 0, st, items, nitems,
                         rowlist, hi, r, full, orig, counts, offsets, o, confirms, 0);
  }
  STQ_HIP(hipGetLastError());
}

void launch_translate(stq_pair *pairs, int64_t count, const uint32_t *orig, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_translate, dim3(blocks_for(count)), dim3(kThreads), 0, st,
                     reinterpret_cast<int2 *>(pairs), count, orig);
  STQ_HIP(hipGetLastError());
}

}  // namespace stqamd
