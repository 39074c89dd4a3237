// plan.hpp -- planning of one 64-row group (device code, shared by the
// one-scene and the batched planners).
//
// The rows of a group are sorted by chain end.  Swept as one body, the group
// costs 64 lane steps per position of [min p + 1, max hi]; when a few rows
// have much longer chains than the rest (heavy-tailed extents), most lanes
// idle through their tails.  So the group is cut after the row k that
// minimises
//     (hi_k - q0 + 1)                                  body: 64 lanes per step
//   + tail_cost * sum_{j > k} (hi_j - hi_k) / 64      tails: one row per lane-step
//   + kTailFixed if k is not the last row              the tail item's own set-up
// (all in wave steps).  The body is swept as before over [q0, hi_k]; the rows above k become a tail
// item over (hi_k, max hi], swept with positions across the lanes and the
// rows taken in turn.  Without long rows the cut is the last row (no tail).
#pragma once

#include <climits>

#include "stq_internal.hpp"
#include "sweep.hpp"

namespace stqamd {

constexpr int kTailFlag = 1 << 30;  // in Item::nrows: a tail item
constexpr int kRowsMask = 127;      // Item::nrows & kRowsMask = rows of the item
constexpr int kCapChunks = 64;      // capped planning: at most this many items per body or tail
constexpr double kTailFixed = 256.0;  // wave steps charged for making a tail item at all

// Most items a capped plan (plan_group with cap) makes for one scene of n
// boxes: each 64-row group's body and tail at most min(kCapChunks, chunks of
// the longest possible span) items.
inline int64_t capped_item_bound(int64_t n, int chunk) {
  if (n < 2) return 0;
  const int64_t spans = (n - 1 + chunk - 1) / chunk;
  return 2 * ((n + 63) / 64) * (spans < kCapChunks ? spans : (int64_t)kCapChunks);
}

// Group record: body {rowoff, nrows, q0, q1}, tail {rowoff, rows, q0, q1}
// (rows = 0: none), chunk sizes {body, tail}.
struct GroupPlan {
  int4 body, tail;
  int2 chunk;
  uint32_t nitems;
  unsigned long long steps;  // lane steps of the body + (row, q) tests of the tail
};

__device__ inline int64_t chunk_count(int64_t span, int ch) {
  return span > 0 ? (span + ch - 1) / ch : 0;
}

// All 64 lanes call this; lane j < nrows holds row j of the group (position p,
// chain end h, non-decreasing in j).  The result is the same in every lane.
// tail_cost <= 0: no tails.
__device__ inline GroupPlan plan_group(int lane, int nrows, int p, int h, int rowoff, int chunk,
                                       bool cap, float tail_cost) {
  const bool act = lane < nrows;
  int pm = act ? p : INT_MAX, hm = act ? h : INT_MIN;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    pm = min(pm, __shfl_xor(pm, o, kWave));
    hm = max(hm, __shfl_xor(hm, o, kWave));
  }
  const int q0 = pm + 1, hmax = hm;
  int kc = nrows - 1;  // last row of the body
  if (tail_cost > 0.f) {
    // inclusive prefix sum of the chain ends
    long long pre = act ? (long long)h : 0;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const long long y = __shfl_up(pre, o, kWave);
      if (lane >= o) pre += y;
    }
    const long long tot = __shfl(pre, kWave - 1, kWave);
    double cost = INFINITY;
    if (act) {
      const double body = (double)max((long long)h - q0 + 1, 0ll);
      const double tail = (double)(tot - pre) - (double)(nrows - 1 - lane) * (double)h;
      cost = body + (double)tail_cost * tail * (1.0 / 64.0) + (lane < nrows - 1 ? kTailFixed : 0.0);
    }
    double cmin = cost;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cmin = fmin(cmin, __shfl_xor(cmin, o, kWave));
    const uint64_t at = __builtin_amdgcn_ballot_w64(act && cost == cmin);
    kc = 63 - __builtin_clzll(at);  // the largest such row: fewest tail rows
  }
  const int hk = __shfl(h, kc, kWave);
  const int R = nrows - 1 - kc;
  GroupPlan g;
  const int q1 = R > 0 ? hk : hmax;
  g.body = make_int4(rowoff, nrows, q0, q1);
  const int64_t sb = (int64_t)q1 - q0 + 1;
  int cb = chunk;
  if (cap && chunk_count(sb, cb) > kCapChunks) cb = (int)((sb + kCapChunks - 1) / kCapChunks);
  int64_t nb = chunk_count(sb, cb);
  unsigned long long steps = sb > 0 ? (unsigned long long)sb * 64ull : 0ull;
  int ct = 64;
  int64_t nt = 0;
  if (R > 0) {
    g.tail = make_int4(rowoff + kc + 1, R, hk + 1, hmax);
    const int64_t st = (int64_t)hmax - hk;
    // a tail item tests about as many (row, q) pairs as a body item of `chunk`
    ct = (int)min<int64_t>(((int64_t)chunk * 64 / R + 63) / 64 * 64, INT_MAX / 2);
    if (cap && chunk_count(st, ct) > kCapChunks) ct = (int)((st + kCapChunks - 1) / kCapChunks);
    nt = chunk_count(st, ct);
    steps += (unsigned long long)R * (unsigned long long)st;
  } else {
    g.tail = make_int4(0, 0, 0, -1);
  }
  g.chunk = make_int2(cb, ct);
  g.nitems = (uint32_t)(nb + nt);
  g.steps = steps;
  return g;
}

// Planning windows: at most kMaxWindow64 rows with 64-bit keys (chain end,
// position), kMaxWindow32 with 32-bit keys (group_key, position): 16 KB of
// LDS either way.
constexpr int kMaxWindow64 = 2048;
constexpr int kMaxWindow32 = 4096;
constexpr int kWinPosBits = 12;  // position bits of a 32-bit window key
// Threads per planning block (one window at a time): 8 waves share the
// window's sort and its 64 groups, and four blocks fit on a CU (all 733
// windows of a 3M-box scene run in one round).
constexpr int kPlanThreads = 512;

// Keys of one planning window into LDS (key[0 .. W)), ending with a barrier:
// with qrec (Q) (group_key, position), else (chain end, position); entries
// past cnt are ~0.  A thread's loads are all in flight at once.
template <bool Q, class K>
__device__ inline void plan_window_keys(const int *__restrict__ hi, const QRec *__restrict__ qrec,
                                        int64_t wbase, int cnt, int W, K *key) {
  constexpr int J = (Q ? kMaxWindow32 : kMaxWindow64) / kPlanThreads;
  uint32_t a[J], b[J];
#pragma unroll
  for (int i = 0; i < J; ++i) {
    const int j = threadIdx.x + i * kPlanThreads;
    a[i] = b[i] = 0u;
    if (j < cnt) {
      if constexpr (Q) {
        const QRec r = qrec[wbase + j];
        a[i] = r.lo;
        b[i] = r.hi;
      } else {
        a[i] = (uint32_t)hi[wbase + j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < J; ++i) {
    const int j = threadIdx.x + i * kPlanThreads;
    if (j < W) {
      K k = ~K(0);
      if (j < cnt) {
        if constexpr (Q) {
          QRec r;
          r.lo = a[i];
          r.hi = b[i];
          k = (K)(((uint64_t)group_key(r) << kWinPosBits) | (uint32_t)j);
        } else {
          k = (K)(((uint64_t)a[i] << 32) | (uint32_t)j);
        }
      }
      key[j] = k;
    }
  }
  __syncthreads();
}

// Groups of one wave whose row loads are in flight together.
constexpr int kPlanBatch = 4;
// Reduced-precision planner, after the window sort (keys (group_key, window
// offset)): the rows of the wave's groups (k = w, w + 8, ...), their
// chain-end loads (and the row features, without tails) issued kPlanBatch
// groups at a time.  Writes rowlist/rowfeat unless tails (the wave sort
// reorders the rows first), and leaves each row's window offset in off16[j]
// and chain end in hend[j] for the group plans.  hend[j] is the LDS word of
// key[j], read before by the same lane.
__device__ inline void plan_rows_q(const int *__restrict__ hi, const uint2 *__restrict__ rowf,
                                   bool tails, const uint32_t *key, uint16_t *off16, int *hend,
                                   int64_t wbase, int cnt, int ng, int w, int lane,
                                   int2 *__restrict__ rowlist, uint2 *__restrict__ rowfeat) {
  for (int k0 = w; k0 < ng; k0 += kPlanBatch * (kPlanThreads / 64)) {
    int o[kPlanBatch], h[kPlanBatch];
    uint2 f[kPlanBatch];
#pragma unroll
    for (int i = 0; i < kPlanBatch; ++i) {
      const int j = 64 * (k0 + i * (kPlanThreads / 64)) + lane;
      o[i] = j < cnt ? (int)(key[j] & ((1u << kWinPosBits) - 1)) : -1;
    }
#pragma unroll
    for (int i = 0; i < kPlanBatch; ++i) {
      h[i] = o[i] >= 0 ? hi[wbase + o[i]] : 0;
      f[i] = o[i] >= 0 && rowf && !tails ? rowf[wbase + o[i]] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int i = 0; i < kPlanBatch; ++i) {
      const int j = 64 * (k0 + i * (kPlanThreads / 64)) + lane;
      if (o[i] >= 0) {
        off16[j] = (uint16_t)o[i];
        hend[j] = h[i];
        if (!tails) {
          rowlist[wbase + j] = make_int2((int)(wbase + o[i]), h[i]);
          if (rowf) rowfeat[wbase + j] = f[i];
        }
      }
    }
  }
}

// Bitonic sort, ascending, of W keys (W a power of two) in LDS by the NT
// threads of the block; ends with a barrier.
template <int NT, class K>
__device__ inline void window_sort(K *key, int W) {
  for (int size = 2; size <= W; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int j = threadIdx.x; j < W; j += NT) {
        const int pj = j ^ stride;
        if (pj > j) {
          const bool up = (j & size) == 0;
          const K a = key[j], c = key[pj];
          if ((a > c) == up) {
            key[j] = c;
            key[pj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// The same sort with each thread holding E = W / NT consecutive keys in
// registers: the compare-exchanges of strides below E stay in registers, the
// others go through LDS (write, barrier, read the partner, barrier): of the
// 78 stages of a 4096-key window, 55 go through LDS with E = 4 and 45 with
// E = 8.  Ends with the sorted keys in LDS and a barrier.
template <int NT, int E, class K>
__device__ inline void window_sort_blocked(K *key) {
  constexpr int W = NT * E;
  K r[E];
  const int base = threadIdx.x * E;
#pragma unroll
  for (int i = 0; i < E; ++i) r[i] = key[base + i];
  for (int size = 2; size <= W; size <<= 1) {
    for (int stride = size >> 1; stride >= E; stride >>= 1) {  // across threads
#pragma unroll
      for (int i = 0; i < E; ++i) key[base + i] = r[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int j = base + i, pj = j ^ stride;
        const K o = key[pj];
        const bool up = (j & size) == 0;
        r[i] = ((j < pj) == up) ? (r[i] < o ? r[i] : o) : (r[i] < o ? o : r[i]);
      }
      __syncthreads();
    }
#pragma unroll
    for (int stride = E / 2; stride > 0; stride >>= 1) {  // within the thread's keys
      if (stride < size) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int pi = i ^ stride;
          if (pi > i) {
            const bool up = ((base + i) & size) == 0;
            const K a = r[i], c = r[pi];
            const bool sw = (a > c) == up;
            r[i] = sw ? c : a;
            r[pi] = sw ? a : c;
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < E; ++i) key[base + i] = r[i];
  __syncthreads();
}

// Sort of a window of W <= WMAX keys (W a power of two): blocked for W >= NT.
template <int NT, int WMAX, class K>
__device__ inline void window_sort_any(K *key, int W) {
  static_assert(WMAX <= 8 * NT, "at most 8 keys per thread");
  if (WMAX >= 8 * NT && W == 8 * NT)
    window_sort_blocked<NT, (WMAX >= 8 * NT ? 8 : 1)>(key);
  else if (WMAX >= 4 * NT && W == 4 * NT)
    window_sort_blocked<NT, (WMAX >= 4 * NT ? 4 : 1)>(key);
  else if (WMAX >= 2 * NT && W == 2 * NT)
    window_sort_blocked<NT, (WMAX >= 2 * NT ? 2 : 1)>(key);
  else if (W == NT)
    window_sort_blocked<NT, 1>(key);
  else
    window_sort<NT>(key, W);
}

// Stable LSD radix sort of the W 32-bit window keys of the reduced-precision
// planner, (group_key << kWinPosBits | position) or all ones past the rows, by
// bits kWinPosBits .. kWinPosBits + 15 in two 8-bit passes (the group key has
// 15 bits and the keys arrive in position order, so stability leaves equal
// group keys in position order: the order of the full keys).  W >= NT, a
// power of two.  In a pass wave w ranks keys [w W/NW, (w+1) W/NW) in rounds
// of 64 (the lanes of equal digit found with 8 ballots, the rank taken from
// the wave's running count of the digit in LDS, the lowest such lane adding
// their number), the counts become each (wave, digit) run's start, and every
// key is written at its place.  cnt: NW x 256 words, dsum: 256 words of LDS.
// About 2.5 times fewer instructions than the bitonic network, and 8 barriers
// instead of a barrier pair per network stage.  Ends with a barrier.
template <int NT>
__device__ inline void window_radix_sort(uint32_t *key, int W, unsigned *cnt, unsigned *dsum) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int E = W / NT;  // 1 .. 8 keys per lane
  const int per = W / NW;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const int sh = kWinPosBits + 8 * pass;
    for (int i = threadIdx.x; i < NW * 256; i += NT) cnt[i] = 0;
    __syncthreads();
    uint32_t k[8];
    unsigned rk[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (r < E) {
        k[r] = key[w * per + r * 64 + lane];
        const unsigned d = (k[r] >> sh) & 255u;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const bool bit = (d >> b) & 1u;
          const uint64_t bb = __builtin_amdgcn_ballot_w64(bit);
          peers &= bit ? bb : ~bb;
        }
        const unsigned before = (unsigned)__popcll(peers & lt);
        const unsigned c0 = cnt[w * 256 + d];
        rk[r] = c0 + before;
        if (before == 0) cnt[w * 256 + d] = c0 + (unsigned)__popcll(peers);
      }
    }
    __syncthreads();
    if (threadIdx.x < 256) {  // each (wave, digit) run's start among the digit's keys
      unsigned t = 0;
#pragma unroll
      for (int v = 0; v < NW; ++v) {
        const unsigned c = cnt[v * 256 + threadIdx.x];
        cnt[v * 256 + threadIdx.x] = t;
        t += c;
      }
      dsum[threadIdx.x] = t;
    }
    __syncthreads();
    if (w == 0) {  // exclusive scan of the digit totals, four per lane
      unsigned v[4], sum = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = dsum[4 * lane + i];
        sum += v[i];
      }
      unsigned inc = sum;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned u = (unsigned)__shfl_up((int)inc, o, 64);
        if (lane >= o) inc += u;
      }
      unsigned run = inc - sum;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        dsum[4 * lane + i] = run;
        run += v[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (r < E) {
        const unsigned d = (k[r] >> sh) & 255u;
        key[dsum[d] + cnt[w * 256 + d] + rk[r]] = k[r];
      }
    }
    __syncthreads();
  }
}

// Sorts the keys of the 64 lanes ascending (a bitonic network over lane
// shuffles); every lane calls this.
__device__ inline uint64_t wave_sort_u64(uint64_t k, int lane) {
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const uint64_t o = __shfl_xor(k, stride, 64);
      const bool up = (lane & size) == 0;
      const bool lower = (lane & stride) == 0;
      k = (lower == up) ? (k < o ? k : o) : (k < o ? o : k);
    }
  }
  return k;
}

// Items of one group at items[o .. o + n): the body's chunks, then the tail's.
__device__ inline void write_group_items(const int4 body, const int4 tail, const int2 ch,
                                         int64_t o, int64_t n, Item *__restrict__ items) {
  const int64_t nb = chunk_count((int64_t)body.w - body.z + 1, ch.x);
  for (int64_t c = 0; c < n; ++c) {
    Item it;
    if (c < nb) {
      it.rowoff = body.x;
      it.nrows = body.y;
      it.q0 = body.z + (int)(c * ch.x);
      it.q1 = (int)min<int64_t>((int64_t)it.q0 + ch.x - 1, body.w);
    } else {
      it.rowoff = tail.x;
      it.nrows = tail.y | kTailFlag;
      it.q0 = tail.z + (int)((c - nb) * ch.y);
      it.q1 = (int)min<int64_t>((int64_t)it.q0 + ch.y - 1, tail.w);
    }
    items[o + c] = it;
  }
}

}  // namespace stqamd
