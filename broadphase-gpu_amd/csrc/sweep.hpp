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
// Reduced-precision form: the two non-sweep axes quantised outward to 14-bit
// levels, packed (axis 1 in the low half, axis 2 in the high half) for the
// lower and the upper bounds, with a kind code in bit 14 of each half (below);
// and the exact FP64 box (sweep axis first) with its ids for the confirmation
// of every candidate the levels let through.
struct alignas(8) QRec {
  uint32_t lo, hi;
};
// Records allocated past the last position of a QRec array: the sweep loads
// every block of 64 positions whole, up to 127 past an item's last one.
constexpr int kQPad = 128;
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
// Adds one box to a thread's axis statistics.
__device__ inline void axis_stats_add(const stq_aabb_f64 &b, double *acc, double *lo, double *up) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double l = b.min[a], u = b.max[a];
    const double c = 0.5 * (l + u);
    acc[4 * a + 0] += l + u;
    acc[4 * a + 1] += l * l + u * u;
    acc[4 * a + 2] += u - l;
    acc[4 * a + 3] += c * c;
    // the quantisation ranges cover the finite bounds: an infinite bound
    // clamps to the end of its range instead of emptying the levels of the scene
    if (isfinite(l)) lo[a] = fmin(lo[a], l);
    if (isfinite(u)) up[a] = fmax(up[a], u);
  }
}
// Adds boxes i, i + step, ... below i1, four loads in flight per thread (the
// reads are latency-bound with one box of 48 bytes per thread at a time).
__device__ inline void axis_stats_range(const stq_aabb_f64 *__restrict__ boxes, int64_t i,
                                        int64_t i1, int64_t step, double *acc, double *lo,
                                        double *up) {
  for (; i + 3 * step < i1; i += 4 * step) {
    stq_aabb_f64 b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = boxes[i + u * step];
#pragma unroll
    for (int u = 0; u < 4; ++u) axis_stats_add(b[u], acc, lo, up);
  }
  for (; i < i1; i += step) axis_stats_add(boxes[i], acc, lo, up);
}
// The same statistics, also writing the sort keys of every axis (reduced
// precision): keys3[a n + i] = image of the rounded-down minimum on axis a;
// *bad is set for a NaN or inverted box.
__device__ inline void axis_stats_keys_range(const stq_aabb_f64 *__restrict__ boxes, int64_t i,
                                             int64_t i1, int64_t step, double *acc, double *lo,
                                             double *up, uint32_t *__restrict__ keys3, int64_t n,
                                             bool &bad) {
  auto put = [&](const stq_aabb_f64 &b, int64_t j) __attribute__((always_inline)) {
    axis_stats_add(b, acc, lo, up);
    bad |= !(b.min[0] <= b.max[0]) || !(b.min[1] <= b.max[1]) || !(b.min[2] <= b.max[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) keys3[a * n + j] = key_of(round_down_f32(b.min[a]));
  };
  for (; i + 3 * step < i1; i += 4 * step) {
    stq_aabb_f64 b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = boxes[i + u * step];
#pragma unroll
    for (int u = 0; u < 4; ++u) put(b[u], i + u * step);
  }
  for (; i < i1; i += step) put(boxes[i], i);
}
// Combines the axis statistics of the block's threads (12 sums, 3 minima,
// 3 maxima; NT threads, all call this) into out[0 .. kAxisStats): shuffle
// trees inside each wave, then the waves in order, so the result is the same
// on every call.
template <int NT>
__device__ inline void block_axis_stats(const double *acc, const double *lo, const double *up,
                                        double *__restrict__ out) {
  __shared__ double wv[NT / 64][kAxisStats];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kAxisStats; ++k) {
    double v = k < 12 ? acc[k] : (k < 15 ? lo[k - 12] : up[k - 15]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(v, o, 64);
      v = k < 12 ? v + y : (k < 15 ? fmin(v, y) : fmax(v, y));
    }
    if (lane == 0) wv[w][k] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < kAxisStats) {
    const int k = threadIdx.x;
    double v = wv[0][k];
    for (int i = 1; i < NT / 64; ++i)
      v = k < 12 ? v + wv[i][k] : (k < 15 ? fmin(v, wv[i][k]) : fmax(v, wv[i][k]));
    out[k] = v;
  }
}

// Outward 14-bit levels: t = (x - origin) * scale is monotone in x under IEEE
// rounding, so floor(t(lo)) <= ceil(t(hi')) whenever lo <= hi': a pair that
// overlaps in FP64 always overlaps in levels (a superset test).
constexpr double kLevelMax = 16383.0;
__device__ inline uint32_t level_down(double x, double o, double sc) {
  if (sc == 0.0) return 0u;
  const double t = fmin(fmax(floor((x - o) * sc), 0.0), kLevelMax);
  return (uint32_t)t;
}
__device__ inline uint32_t level_up(double x, double o, double sc) {
  if (sc == 0.0) return (uint32_t)kLevelMax;
  const double t = fmin(fmax(ceil((x - o) * sc), 0.0), kLevelMax);
  return (uint32_t)t;
}

// Kind codes.  The sweep tests a row p against a column q (p sorted before q)
// by packed 16-bit differences of (level + 16384 * code) per half: equal codes
// leave the level test as it is, and unequal codes make one of the two
// differences negative whatever the levels (levels are in [0, 16383]).  A box
// is stored in its column form; a row takes its row form.  Codes per half
// (axis 1, axis 2):
//   column: face (0,0), vertex (1,0), edge (0,1)
//   row:    vertex (0,0), face (1,0), edge (0,1)
// so row and column codes are equal exactly for vertex-face, face-vertex and
// edge-edge, the pairs the type filter keeps (src/aabb.cu:12-40): the level
// test drops the other kind pairs at no cost.  No column code is (1,1), so in
// every half a column's upper level is below 0x7fff or its lower level above
// 0: the empty row {lo 0x7fff7fff, hi 0} meets no column.
constexpr uint32_t kCodeBits = 0x40004000u;
__host__ __device__ inline uint32_t column_code(int typebit) {
  return typebit == 1 ? 0x00004000u : (typebit == 2 ? 0x40000000u : 0u);
}
__host__ __device__ inline uint32_t row_code(int typebit) {
  return typebit == 4 ? 0x00004000u : (typebit == 2 ? 0x40000000u : 0u);
}
// Row form of a stored (column form) record.
__device__ __forceinline__ QRec row_form(QRec c) {
  const uint32_t k = c.lo & kCodeBits;
  const int t = k == 0x00004000u ? 1 : (k == 0x40000000u ? 2 : 4);
  c.lo = (c.lo & ~kCodeBits) | row_code(t);
  c.hi = (c.hi & ~kCodeBits) | row_code(t);
  return c;
}
// Grouping key of a row for the reduced-precision sweep: the centre of its
// axis-1 levels (twice the centre, kind code dropped).
__device__ __forceinline__ uint32_t group_key(QRec r) {
  return (r.lo & 0x3fffu) + (r.hi & 0x3fffu);
}
// Column form of a box: levels of the two non-sweep axes and the kind code.
__device__ inline QRec make_qrec(double lo1, double hi1, double lo2, double hi2, int typebit,
                                 double o1, double s1, double o2, double s2) {
  const uint32_t code = column_code(typebit);
  QRec q;
  q.lo = (level_down(lo1, o1, s1) | (level_down(lo2, o2, s2) << 16)) | code;
  q.hi = (level_up(hi1, o1, s1) | (level_up(hi2, o2, s2) << 16)) | code;
  return q;
}

// ---- column statistics of position blocks (learned sizing features) ----
// Per block of kColBlock consecutive sorted positions: sums of the log-extent
// (log(extent)/10) and of the centre's place in the scene's level range (0..1)
// on the two non-sweep axes, of their squares, and the kind counts.  An item's
// column features are the statistics of the blocks its position range meets.
constexpr int kColBlock = 32;
struct alignas(16) ColStats {
  float le1, le1sq, le2, le2sq;
  float c1, c1sq, c2, c2sq;
  float nv, ne, nf, cnt;
};
__device__ __forceinline__ float log_extent(double lo, double hi) {
  return logf(fmaxf((float)(hi - lo), 1e-30f)) * 0.1f;
}
__device__ __forceinline__ float centre_place(double lo, double hi, double origin, double norm) {
  return (float)((0.5 * (lo + hi) - origin) * norm);
}

// The row features of a box for learned sizing, packed in 8 bytes: the
// log-extents of both non-sweep axes as FP16, the centre places (0..1) as
// 14- and 16-bit fractions and the kind (0 vertex, 1 edge, 2 face) in the two
// top bits of the first.  The gather writes them by sorted position and the
// planner copies them into group order, so the features of a group's rows
// are one contiguous 512-byte read instead of 64 scattered FP64 records.
__device__ inline uint2 row_feat(double lo1, double hi1, double lo2, double hi2, int type,
                                 double o1, double n1, double o2, double n2) {
  const _Float16 l1 = (_Float16)log_extent(lo1, hi1), l2 = (_Float16)log_extent(lo2, hi2);
  const float c1 = fminf(fmaxf(centre_place(lo1, hi1, o1, n1), 0.f), 1.f);
  const float c2 = fminf(fmaxf(centre_place(lo2, hi2, o2, n2), 0.f), 1.f);
  const uint32_t kind = type == 1 ? 0u : (type == 2 ? 1u : 2u);
  uint2 f;
  f.x = (uint32_t)__builtin_bit_cast(uint16_t, l1) |
        ((uint32_t)__builtin_bit_cast(uint16_t, l2) << 16);
  f.y = (uint32_t)(c1 * 16383.f + 0.5f) | (kind << 14) | ((uint32_t)(c2 * 65535.f + 0.5f) << 16);
  return f;
}
struct RowFeat {
  float le1, le2, c1, c2;
  int kind;  // 0 vertex, 1 edge, 2 face
};
__device__ inline RowFeat row_feat_unpack(uint2 f) {
  RowFeat r;
  r.le1 = (float)__builtin_bit_cast(_Float16, (uint16_t)(f.x & 0xffffu));
  r.le2 = (float)__builtin_bit_cast(_Float16, (uint16_t)(f.x >> 16));
  r.c1 = (float)(f.y & 0x3fffu) * (1.f / 16383.f);
  r.kind = (int)((f.y >> 14) & 3u);
  r.c2 = (float)(f.y >> 16) * (1.f / 65535.f);
  return r;
}
// Adds sorted position s to its block.  Every lane of the wave calls this,
// with valid = false past the last position; the wave's lanes hold 64
// consecutive positions from a multiple of 64, so a block is 32 consecutive
// lanes.  norm = scale / kLevelMax of the scene.
__device__ inline void col_stats_put(bool valid, int64_t s, double lo1, double hi1, double lo2,
                                     double hi2, int type, double o1, double n1, double o2,
                                     double n2, ColStats *__restrict__ out) {
  float v[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = 0.f;
  if (valid) {
    const float le1 = log_extent(lo1, hi1), le2 = log_extent(lo2, hi2);
    const float c1 = centre_place(lo1, hi1, o1, n1), c2 = centre_place(lo2, hi2, o2, n2);
    v[0] = le1;
    v[1] = le1 * le1;
    v[2] = le2;
    v[3] = le2 * le2;
    v[4] = c1;
    v[5] = c1 * c1;
    v[6] = c2;
    v[7] = c2 * c2;
    v[8] = type == 1 ? 1.f : 0.f;
    v[9] = type == 2 ? 1.f : 0.f;
    v[10] = type == 4 ? 1.f : 0.f;
    v[11] = 1.f;
  }
#pragma unroll
  for (int k = 0; k < 12; ++k)
#pragma unroll
    for (int o = kColBlock / 2; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  if (valid && (threadIdx.x & (kColBlock - 1)) == 0) {
    ColStats c;
    c.le1 = v[0];
    c.le1sq = v[1];
    c.le2 = v[2];
    c.le2sq = v[3];
    c.c1 = v[4];
    c.c1sq = v[5];
    c.c2 = v[6];
    c.c2sq = v[7];
    c.nv = v[8];
    c.ne = v[9];
    c.nf = v[10];
    c.cnt = v[11];
    out[s / kColBlock] = c;
  }
}

// Per-scene sweep axis and level ranges; for the reduced-precision form also
// the quantisation of the sweep axis's keys (below).
struct SceneParams {
  int axis;
  uint32_t kmax;   // largest sweep key: 2^bits - 1
  QParams qp;
  double korigin;  // sweep-axis coordinate of key 0
  double kscale;   // sweep keys per unit of the sweep axis (0: every key equal)
};

// Sweep keys of the reduced-precision form.  The FP32 outward bounds of the
// sweep axis (the reference's float rounding) are quantised outward once more,
// to integers in [0, kmax] over the scene's key range (set_sweep_keys):
//   down(f) = floor((f - korigin) * kscale),  up(f) = ceil((f - korigin) * kscale),
// clamped.  Both are non-decreasing in f, so a pair that overlaps in FP64
// overlaps in keys (min_q <= max_p gives down(min_q) <= up(max_p)), and the
// sort, the chain ends and the planning run on these integers; the FP64
// confirmation tests the sweep axis again.  The sort key of a box is exactly
// its down(min), so the keys in sorted order are non-decreasing.
__device__ inline uint32_t skey_down(float f, const SceneParams &p) {
  if (p.kscale == 0.0) return 0u;
  const double t = floor(((double)f - p.korigin) * p.kscale);
  return (uint32_t)fmin(fmax(t, 0.0), (double)p.kmax);
}
__device__ inline uint32_t skey_up(float f, const SceneParams &p) {
  if (p.kscale == 0.0) return p.kmax;
  const double t = ceil(((double)f - p.korigin) * p.kscale);
  return (uint32_t)fmin(fmax(t, 0.0), (double)p.kmax);
}
// Inverse of key_of(float): the float whose order-preserving image is k.
__device__ inline float float_of_key(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __builtin_bit_cast(float, u);
}
// Bits of the sweep keys for a scene of at most n boxes: 16 up to 2^22 boxes
// (two sort passes; a uniform scene puts at most 64 boxes in a key, which
// lengthens a chain by about as many positions: 1% on the 3M-box cloth), 24
// up to 2^28, else 32.
inline int sweep_key_bits(int64_t n) { return n <= (1 << 22) ? 16 : (n <= (1 << 28) ? 24 : 32); }
// The quantisation of the chosen axis over the scene's finite range [lo, up]
// on that axis; launch_sweep_range then narrows it to the bulk of the scene.
__device__ inline void set_sweep_keys(SceneParams &p, double lo, double up, int bits) {
  p.kmax = bits >= 32 ? 0xffffffffu : (1u << bits) - 1u;
  const double range = up - lo;
  const double sc = (range > 0 && isfinite(range)) ? (double)p.kmax / range : 0.0;
  p.korigin = isfinite(lo) ? lo : 0.0;
  p.kscale = isfinite(sc) ? sc : 0.0;
}

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
// dependent load rounds instead of the 2 log_2 L of a binary search.  The
// search is bound by load throughput more than by latency: on the 512-scene
// batch K = 3 beat 1, 2, 4, 8 and 16 (1.0 ms; 16 took 1.6 ms).
template <class T, int K = 3>
__device__ __forceinline__ int64_t last_le(const T *__restrict__ smin, int64_t s, int64_t end,
                                           T x) {
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

// Chain ends of a tile of consecutive sorted rows, the minima staged in LDS.
// A block of kCeThreads threads takes the rows [s0, s0 + kCeRows) and copies
// the minima of positions [s0, w1) into LDS, w1 = min(n, s0 + kCeWin): a
// tile's rows see at least 3 kCeRows positions ahead in LDS.  Each thread
// then runs the binary searches of its rows together (4 float or 2 double
// rows, their LDS reads in flight at once): the last q in [s, min(end, w1))
// with smin[q] <= smax[s].  A row whose answer is the last staged position
// while its scene goes on continues with last_le in memory.  end_of(s) is
// the end of the row's scene; rows outside [r0, r1) get empty chains
// (hi = s).  Returns the sum of the thread's chain lengths.
constexpr int kCeThreads = 1024;
template <class T>
constexpr int kCeWin = 65536 / (int)sizeof(T);  // 64 KB of LDS
template <class T>
constexpr int kCeRows = kCeWin<T> / 4;
template <class T, class EndOf>
__device__ inline unsigned long long chain_end_tile(const T *__restrict__ smin,
                                                    const T *__restrict__ smax, int64_t n,
                                                    int64_t s0, int64_t r0, int64_t r1, T *lds,
                                                    EndOf end_of, int *__restrict__ hi) {
  constexpr int R = kCeRows<T> / kCeThreads;  // 4 (float) or 2 (double)
  static_assert(R >= 1 && kCeRows<T> == R * kCeThreads, "whole rows per thread");
  const int64_t w1 = s0 + kCeWin<T> < n ? s0 + kCeWin<T> : n;
  const int nw = (int)(w1 - s0);
  // the rows' own loads first: they stay in flight with the window's
  int lo[R], up[R];
  int64_t end[R];
  T x[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t s = s0 + threadIdx.x + (int64_t)r * kCeThreads;
    lo[r] = 0;
    up[r] = 1;  // done: lds[0] exists (nw >= 1)
    end[r] = s;
    x[r] = T(0);
    if (s < n && s >= r0 && s < r1) {
      end[r] = end_of(s);
      lo[r] = (int)(s - s0);
      up[r] = (int)((end[r] < w1 ? end[r] : w1) - s0);
      x[r] = smax[s];
    }
  }
  __syncthreads();  // the previous tile's searches are done with the LDS
  if (nw == kCeWin<T>) {
    // a whole window in 16-byte loads, all of a thread's in flight at once
    // (s0 is a multiple of kCeRows entries, so they are aligned)
    constexpr int L = kCeWin<T> * (int)sizeof(T) / 16 / kCeThreads;
    static_assert(L * 16 * kCeThreads == kCeWin<T> * (int)sizeof(T), "whole loads per thread");
    const uint4 *src = reinterpret_cast<const uint4 *>(smin + s0);
    uint4 v[L];
#pragma unroll
    for (int l = 0; l < L; ++l) v[l] = src[threadIdx.x + l * kCeThreads];
#pragma unroll
    for (int l = 0; l < L; ++l) reinterpret_cast<uint4 *>(lds)[threadIdx.x + l * kCeThreads] = v[l];
  } else {
    for (int i = threadIdx.x; i < nw; i += kCeThreads) lds[i] = smin[s0 + i];
  }
  __syncthreads();
  // invariant: lds[lo] <= x, the answer lies in [lo, up)
  while (true) {
    T v[R];
    int mid[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      mid[r] = (lo[r] + up[r]) >> 1;
      v[r] = lds[mid[r]];
    }
    bool more = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (up[r] - lo[r] > 1) {
        if (v[r] <= x[r])
          lo[r] = mid[r];
        else
          up[r] = mid[r];
      }
      more |= up[r] - lo[r] > 1;
    }
    if (!more) break;
  }
  unsigned long long len = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t s = s0 + threadIdx.x + (int64_t)r * kCeThreads;
    if (s >= n) continue;
    if (s < r0 || s >= r1) {
      hi[s] = (int)s;
      continue;
    }
    int64_t q = s0 + lo[r];
    if (q == w1 - 1 && w1 < end[r]) q = last_le(smin, q, end[r], x[r]);
    hi[s] = (int)q;
    len += (unsigned long long)(q - s);
  }
  return len;
}

// keys3 (reduced precision; null: none): the pass also writes the sort keys
// of all three axes (3 n words) and sets kErrInverted in *err for a NaN or
// inverted box.
void launch_axis_stats(const stq_aabb_f64 *boxes, int64_t n, double *partial, int nblocks,
                       hipStream_t st, uint32_t *keys3 = nullptr, unsigned *err = nullptr);
// The scene's sweep axis (axis policy 0..4), level ranges and the
// quantisation of its sweep keys (sweep_key_bits(n) bits) from the statistics
// of nblocks blocks, into *out on the device.
//
// launch_sweep_range narrows each scene's sweep-key range to its bulk: a
// sample of at most 512 boxes of the scene, the 0.5% quantile of their
// minima and the 99.5% quantile of their maxima on the sweep axis, widened
// by an eighth of their distance on each side (never past the finite range).  A few far-away boxes then no
// longer spread the keys of all the others over a range they barely use
// (they share the end keys, which clamping keeps outward).  seg_off null:
// one scene of n boxes.
void launch_sweep_range(const stq_aabb_f64 *boxes, const int64_t *seg_off, int nscenes, int64_t n,
                        SceneParams *sp, hipStream_t st);
void launch_axis_final(const double *partial, int nblocks, int64_t n, int policy,
                       SceneParams *out, hipStream_t st);
// The sweep axis is sp->axis where sp is given, else axis.  The FP32 form
// needs sp (its level range).
void launch_keys(bool f32, const stq_aabb_f64 *boxes, int64_t n, int axis, const SceneParams *sp,
                 void *keys, uint32_t *vals, unsigned *err, hipStream_t st);
// cols: column statistics of every kColBlock positions, rowf: the row
// features of every position (row_feat); null: not made.
void launch_gather(bool f32, const stq_aabb_f64 *boxes, const uint32_t *order, int64_t n,
                   int axis, const SceneParams *sp, void *smin, void *smax, void *rec, QRec *qrec,
                   ColStats *cols, uint2 *rowf, hipStream_t st);
// Rows outside [r0, r1) (r1 < 0: n) get empty chains.
void launch_chain_end(bool f32, const void *smin, const void *smax, int64_t n, int *hi,
                      unsigned long long *cand, hipStream_t st, int64_t r0 = 0, int64_t r1 = -1);
// tail_cost: weight of tail tests in the group cut (plan.hpp); 0 = no tails.
// qrec (reduced precision): windows are ordered by group_key, so a group's
// rows are close on axis 1 (with tails each group's rows then by chain end); null:
// windows ordered by chain end.  cap: at most kCapChunks items per body and
// per tail (chunks grow instead), so a scene has at most
// capped_item_bound(n, chunk) items.
// rowf (learned sizing; null: none): the row features by sorted position,
// copied into rowfeat in the order of rowlist.
void launch_plan(const int *hi, const QRec *qrec, int64_t n, int W, int chunk, float tail_cost,
                 int2 *rowlist, int4 *groups, int4 *tgroups, int2 *gchunk, uint32_t *nchunks,
                 unsigned long long *lane_steps, hipStream_t st, bool cap = false,
                 const uint2 *rowf = nullptr, uint2 *rowfeat = nullptr);
// Items of every group from the group records (body chunks, then tail chunks).
void launch_items(const int4 *groups, const int4 *tgroups, const int2 *gchunk, const int64_t *ioff,
                  int64_t ngroups, Item *items, hipStream_t st);
enum SweepMode { kCount = 0, kFill = 1, kLearned = 2 };

// Words between the per-XCD item-claim counters of one sweep launch.
constexpr int kClaimStride = 64;

struct SweepArgs {
  const Item *items = nullptr;
  int64_t nitems = 0;            // work items (or selected items when sel != null)
  const int *sel = nullptr;      // optional item selection
  const int2 *rowlist = nullptr;  // the rows in group order: {sorted position, chain end}
  const uint2 *rowfeat = nullptr;  // learned sizing: their row features (row_feat), same order
  const void *rec = nullptr;     // Rec64 (FP64 form) or Exact64 by sorted position
  const QRec *qrec = nullptr;    // reduced-precision form; with Rec64 records: the grouped FP64 form
  const uint32_t *orig = nullptr;
  uint32_t *counts = nullptr;    // kCount, kLearned: exact pairs per item
  const int64_t *offsets = nullptr;  // kFill: output offset; kLearned: scratch offset
  const uint32_t *caps = nullptr;    // kLearned: capacity per item
  int2 *out = nullptr;           // kFill: output; kLearned: scratch
  // kCount, kLearned: FP64 confirmations, one counter per XCD slice
  // kClaimStride / 2 words apart (null: none)
  unsigned long long *confirms = nullptr;
  int zero = 0;                  // always 0 (an opaque runtime zero)
  int xcd_split = 1;             // 1: each XCD sweeps one contiguous slice of items
  // device-resident sizes (batched pipeline): the item count, and for kFill
  // the pair total with the output capacity (nothing is written when the
  // total exceeds it)
  const int64_t *nitems_dev = nullptr;
  const int64_t *total_dev = nullptr;
  int64_t capacity = 0;
  // Zeroed claim counters, one per XCD slice, kClaimStride words apart: each
  // wave takes its first item by its index in the slice and claims the
  // following ones; null: a fixed stride
  unsigned *work = nullptr;
  // With xcd_split: kXcds + 1 item boundaries of the XCD slices (null: equal
  // numbers of items per slice)
  const int64_t *slices = nullptr;
};

// Pairs come out as canonical (min, max) pairs of input ids (orig[] of both
// sorted positions), in the emission order of the sweep.
void launch_sweep(bool f32, int mode, int batch, int nblocks, const SweepArgs &args,
                  hipStream_t st);

// Work split by predicted cost.  Each item's cost is estimated from its shape
// and its expected pairs (before a learned pass the network's prediction,
// log(1 + pairs), bounded by the item's candidates cand[]; before a fill pass
// the exact counts; none before a count pass); the XCD slices then get equal
// shares of the total, in item order.  cost[] is n (or *n_dev) words;
// slices[] gets kXcds + 1 boundaries from the exclusive prefix sums.
void launch_item_cost(const Item *items, int64_t n, const int64_t *n_dev, bool grouped,
                      const float *pred, const uint32_t *cand, const uint32_t *counts,
                      uint32_t *cost, hipStream_t st);
void launch_xcd_slices(const int64_t *prefix, int64_t n, const int64_t *n_dev, int64_t *slices,
                       hipStream_t st);

}  // namespace stqamd
