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

__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }
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
                                                         Geo64 *__restrict__ geo,
                                                         int4 *__restrict__ meta) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (s >= n) return;
  const stq_aabb_f64 &b = boxes[order[s]];
  int a1, a2;
  other_axes(axis, a1, a2);
  smin[s] = b.min[axis];
  smax[s] = b.max[axis];
  Geo64 g;
  g.lo1 = b.min[a1];
  g.hi1 = b.max[a1];
  g.lo2 = b.min[a2];
  g.hi2 = b.max[a2];
  geo[s] = g;
  const int v0 = b.vertex_ids[0], v1 = b.vertex_ids[1], v2 = b.vertex_ids[2];
  meta[s] = make_int4(v0, v1, v2, type_bit(v1, v2));
}

__global__ __launch_bounds__(kThreads) void k_gather_f32(const stq_aabb_f64 *__restrict__ boxes,
                                                         const uint32_t *__restrict__ order,
                                                         int64_t n, int axis,
                                                         float *__restrict__ smin,
                                                         float *__restrict__ smax,
                                                         Geo32 *__restrict__ geo,
                                                         int4 *__restrict__ meta,
                                                         Full64 *__restrict__ full) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (s >= n) return;
  const stq_aabb_f64 &b = boxes[order[s]];
  int a1, a2;
  other_axes(axis, a1, a2);
  smin[s] = round_down_f32(b.min[axis]);
  smax[s] = round_up_f32(b.max[axis]);
  Geo32 g;
  g.lo1 = round_down_f32(b.min[a1]);
  g.hi1 = round_up_f32(b.max[a1]);
  g.lo2 = round_down_f32(b.min[a2]);
  g.hi2 = round_up_f32(b.max[a2]);
  geo[s] = g;
  Full64 f;
  f.lo0 = b.min[axis];
  f.hi0 = b.max[axis];
  f.lo1 = b.min[a1];
  f.hi1 = b.max[a1];
  f.lo2 = b.min[a2];
  f.hi2 = b.max[a2];
  full[s] = f;
  const int v0 = b.vertex_ids[0], v1 = b.vertex_ids[1], v2 = b.vertex_ids[2];
  meta[s] = make_int4(v0, v1, v2, type_bit(v1, v2));
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
// All tests are written with bitwise & and | so that they compile to v_cmp +
// s_and on lane masks instead of short-circuit branches.
__device__ inline bool covertex(int a0, int a1, int a2, const int4 &b) {
  return (a0 == b.x) | (a0 == b.y) | (a0 == b.z) | (a1 == b.x) | (a1 == b.y) | (a1 == b.z) |
         (a2 == b.x) | (a2 == b.y) | (a2 == b.z);
}

template <class G>
__device__ inline bool overlap2(const G &m, const G &q) {
  return (m.hi1 >= q.lo1) & (m.lo1 <= q.hi1) & (m.hi2 >= q.lo2) & (m.lo2 <= q.hi2);
}

__device__ inline bool confirm64(const Full64 &m, const Full64 &q) {
  return (m.hi0 >= q.lo0) & (m.lo0 <= q.hi0) & (m.hi1 >= q.lo1) & (m.lo1 <= q.hi1) &
         (m.hi2 >= q.lo2) & (m.lo2 <= q.hi2);
}

__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Per-wave state of one work item: the 64 rows' records live in VGPRs.
template <class G, bool CONFIRM, bool FILL>
struct Rows {
  int lane, p, myhi;
  G mine;
  int4 mv;
  Full64 mf;
  uint32_t op;
  uint64_t act, accV, accE, accF;
  unsigned cnt;
  int64_t base;
  unsigned long long nconf;

  // Pairs of the rows in mask m (geometric test passed) with sorted q.
  __device__ inline void hit(int q, uint64_t m, const int4 *__restrict__ meta,
                             const Full64 *__restrict__ full, const uint32_t *__restrict__ orig,
                             stq_pair *__restrict__ out) {
    const int4 mq = meta[q];
    m &= (mq.w == 1) ? accV : ((mq.w == 2) ? accE : accF);
    if (m == 0) return;
    bool h = (((m >> lane) & 1ull) != 0) & !covertex(mv.x, mv.y, mv.z, mq);
    if (CONFIRM) {
      const uint64_t hm0 = ballot(h);
      if (hm0 == 0) return;
      nconf += (unsigned long long)__popcll(hm0);
      const Full64 fq = full[q];
      h = h & confirm64(mf, fq);
    }
    const uint64_t hm = ballot(h);
    if (hm == 0) return;
    if (FILL) {
      if (h) {
        const uint32_t oq = orig[q];
        stq_pair pr;
        pr.first = (int)min(op, oq);
        pr.second = (int)max(op, oq);
        out[base + mbcnt64(hm)] = pr;
      }
      base += __popcll(hm);
    } else {
      cnt += (unsigned)__popcll(hm);
    }
  }

  // q in [qa, qb]; CHECK adds the per-lane chain range test p < q <= hi_p.
  template <bool CHECK>
  __device__ inline void walk(int qa, int qb, const G *__restrict__ geo,
                              const int4 *__restrict__ meta, const Full64 *__restrict__ full,
                              const uint32_t *__restrict__ orig, stq_pair *__restrict__ out) {
    int q = qa;
    for (; q + 3 <= qb; q += 4) {
      const G g0 = geo[q], g1 = geo[q + 1], g2 = geo[q + 2], g3 = geo[q + 3];
      bool c0 = overlap2(mine, g0), c1 = overlap2(mine, g1), c2 = overlap2(mine, g2),
           c3 = overlap2(mine, g3);
      if (CHECK) {
        c0 = c0 & (q > p) & (q <= myhi);
        c1 = c1 & (q + 1 > p) & (q + 1 <= myhi);
        c2 = c2 & (q + 2 > p) & (q + 2 <= myhi);
        c3 = c3 & (q + 3 > p) & (q + 3 <= myhi);
      }
      const uint64_t m0 = ballot(c0) & act, m1 = ballot(c1) & act, m2 = ballot(c2) & act,
                     m3 = ballot(c3) & act;
      if ((m0 | m1 | m2 | m3) != 0) {
        if (m0) hit(q, m0, meta, full, orig, out);
        if (m1) hit(q + 1, m1, meta, full, orig, out);
        if (m2) hit(q + 2, m2, meta, full, orig, out);
        if (m3) hit(q + 3, m3, meta, full, orig, out);
      }
    }
    for (; q <= qb; ++q) {
      const G g = geo[q];
      bool c = overlap2(mine, g);
      if (CHECK) c = c & (q > p) & (q <= myhi);
      const uint64_t m = ballot(c) & act;
      if (m) hit(q, m, meta, full, orig, out);
    }
  }
};

// FILL = false: per-item pair counts; FILL = true: write pairs at offsets[it].
// CONFIRM: G holds FP32 outward bounds, every FP32 hit is re-tested on Full64.
template <class G, bool CONFIRM, bool FILL>
__global__ __launch_bounds__(kThreads) void k_sweep(
    const Item *__restrict__ items, int64_t nitems, const int *__restrict__ rowlist,
    const int *__restrict__ hi, const G *__restrict__ geo, const int4 *__restrict__ meta,
    const Full64 *__restrict__ full, const uint32_t *__restrict__ orig,
    uint32_t *__restrict__ counts, const int64_t *__restrict__ offsets,
    stq_pair *__restrict__ out, unsigned long long *__restrict__ confirms) {
  Rows<G, CONFIRM, FILL> r;
  r.lane = threadIdx.x & 63;
  const int wave = uniform((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  const int nwaves = (int)gridDim.x * (kThreads / 64);
  r.nconf = 0;
  for (int64_t it = wave; it < nitems; it += nwaves) {
    const Item item = items[it];
    const bool act = r.lane < item.nrows;
    r.p = act ? rowlist[item.rowoff + r.lane] : INT_MAX;
    r.myhi = act ? hi[r.p] : INT_MIN;
    r.mv = make_int4(0, 0, 0, 0);
    r.op = 0;
    if (act) {
      r.mine = geo[r.p];
      r.mv = meta[r.p];
      if (CONFIRM) r.mf = full[r.p];
      if (FILL) r.op = orig[r.p];
    } else {
      r.mine = G{};
      if (CONFIRM) r.mf = Full64{};
    }
    const int acc = act ? accept_bits(r.mv.w) : 0;
    r.act = ballot(act);
    r.accV = ballot(acc & 1);
    r.accE = ballot(acc & 2);
    r.accF = ballot(acc & 4);
    r.cnt = 0;
    r.base = FILL ? offsets[it] : 0;
    // rows with p < q <= hi_p for every active lane: q in (pmax, hmin]
    int pmax = act ? r.p : INT_MIN, hmin = act ? r.myhi : INT_MAX;
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
      if (q0 < b0) r.template walk<true>(q0, b0 - 1, geo, meta, full, orig, out);
      r.template walk<false>(b0, b1, geo, meta, full, orig, out);
      if (b1 < q1) r.template walk<true>(b1 + 1, q1, geo, meta, full, orig, out);
    } else {
      r.template walk<true>(q0, q1, geo, meta, full, orig, out);
    }
    if (!FILL && r.lane == 0) counts[it] = r.cnt;
  }
  if (CONFIRM && !FILL && r.lane == 0 && r.nconf) atomicAdd(confirms, r.nconf);
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
                   int axis, void *smin, void *smax, void *geo, int4 *meta, Full64 *full,
                   hipStream_t st) {
  if (n <= 0) return;
  if (f32)
    hipLaunchKernelGGL(k_gather_f32, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, order,
                       n, axis, (float *)smin, (float *)smax, (Geo32 *)geo, meta, full);
  else
    hipLaunchKernelGGL(k_gather_f64, dim3(blocks_for(n)), dim3(kThreads), 0, st, boxes, order,
                       n, axis, (double *)smin, (double *)smax, (Geo64 *)geo, meta);
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
                  const int *rowlist, const int *hi, const void *geo, const int4 *meta,
                  const Full64 *full, const uint32_t *orig, uint32_t *counts,
                  const int64_t *offsets, stq_pair *out, unsigned long long *confirms,
                  hipStream_t st) {
  if (nitems <= 0) return;
  const dim3 grid(nblocks), block(kThreads);
  if (f32) {
    if (fill)
      hipLaunchKernelGGL((k_sweep<Geo32, true, true>), grid, block, 0, st, items, nitems, rowlist,
                         hi, (const Geo32 *)geo, meta, full, orig, counts, offsets, out, confirms);
    else
      hipLaunchKernelGGL((k_sweep<Geo32, true, false>), grid, block, 0, st, items, nitems,
                         rowlist, hi, (const Geo32 *)geo, meta, full, orig, counts, offsets, out,
                         confirms);
  } else {
    if (fill)
      hipLaunchKernelGGL((k_sweep<Geo64, false, true>), grid, block, 0, st, items, nitems,
                         rowlist, hi, (const Geo64 *)geo, meta, full, orig, counts, offsets, out,
                         confirms);
    else
      hipLaunchKernelGGL((k_sweep<Geo64, false, false>), grid, block, 0, st, items, nitems,
                         rowlist, hi, (const Geo64 *)geo, meta, full, orig, counts, offsets, out,
                         confirms);
  }
  STQ_HIP(hipGetLastError());
}

}  // namespace stqamd
