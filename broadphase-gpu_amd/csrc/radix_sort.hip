// radix_sort.hip -- stable LSD radix sort of (key, uint32 value) pairs and
// exclusive prefix sums, written for gfx950 (wave64, 160 KB LDS per CU).
//
// Replaces the reference's thrust::sort_by_key(d_sm, d_sm + N, d_mini,
// sort_aabb_x()) (src/simulation.cu:645-649), a comparator merge sort that
// moves the 48-byte MiniBox with every key.  Here the keys are order-
// preserving integer images of the sweep-axis minimum and the values are
// 4-byte input positions; the box records are gathered once after the sort.
//
// One pass per 8-bit digit:
//   k_count   : per-tile digit histogram -> counts[digit * ntiles + tile]
//   scan      : exclusive prefix over counts (digit-major) -> global offsets
//   k_scatter : per-wave ballot multisplit gives each key its stable rank in
//               the tile; keys are staged in LDS in digit order and written
//               out in runs, so consecutive lanes store consecutive addresses.
// A digit that is equal for all keys (one bucket holds all n) is skipped; the
// per-digit histograms of all passes come from one pre-pass.
#include "stq_internal.hpp"
#include "sweep.hpp"

namespace stqamd {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kTile = kSortThreads * kSortItems;  // 4096 keys per tile
constexpr int kWavesPerBlock = kSortThreads / kWave;

__device__ inline uint64_t lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ---- histogram of every digit (pre-pass) ----------------------------------
template <class K>
__global__ __launch_bounds__(kSortThreads) void k_hist_all(const K *__restrict__ keys,
                                                            int64_t n, int ndigits,
                                                            unsigned *__restrict__ hist) {
  __shared__ unsigned h[8][256];
  for (int i = threadIdx.x; i < 8 * 256; i += kSortThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kSortThreads;
  for (int64_t i = (int64_t)blockIdx.x * kSortThreads + threadIdx.x; i < n; i += stride) {
    K k = keys[i];
    for (int d = 0; d < ndigits; ++d) atomicAdd(&h[d][(unsigned)(k >> (8 * d)) & 255u], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ndigits * 256; i += kSortThreads) {
    unsigned v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

// ---- per-tile digit counts --------------------------------------------------
// Tile t covers keys [t * kTile, min(n, (t + 1) * kTile)) and its count of
// digit d goes to counts[d * ntiles + t]; with a segment table (SEG) tile t is
// tiles[t] and its count goes to counts[cbase + d * stride], so one exclusive
// scan over the counts gives every (tile, digit) its place inside the tile's
// own segment (segment-major, then digit, then tile).
__device__ __forceinline__ void tile_of(const SegTile *tiles, int64_t n, int64_t &base, int &cnt) {
  if (tiles) {
    const SegTile t = tiles[blockIdx.x];
    base = t.start;
    cnt = t.count;
  } else {
    base = (int64_t)blockIdx.x * kTile;
    const int64_t rem = n - base;
    cnt = rem < kTile ? (int)rem : kTile;
  }
}
__device__ __forceinline__ int64_t count_index(const SegTile *tiles, int ntiles, int d) {
  if (tiles) {
    const SegTile t = tiles[blockIdx.x];
    return t.cbase + (int64_t)d * t.stride;
  }
  return (int64_t)d * ntiles + blockIdx.x;
}

// With scene parameters (the first pass over the FP32 keys of all three axes)
// the tile reads the keys of its scene's sweep axis, keys + axis * n, and
// sorts their sweep keys: each key is turned into skey_down of its float.
struct KeySource {
  int64_t offset = 0;
  const SceneParams *p = nullptr;
};
__device__ __forceinline__ KeySource key_source(const SegTile *tiles, const SceneParams *sp,
                                                int64_t n) {
  KeySource k;
  if (!sp) return k;
  k.p = sp + (tiles ? tiles[blockIdx.x].scene : 0);
  k.offset = (int64_t)k.p->axis * n;
  return k;
}
template <class K>
__device__ __forceinline__ K source_key(K k, const SceneParams *p) {
  if constexpr (sizeof(K) == 4) {
    if (p) return (K)skey_down(float_of_key((uint32_t)k), *p);
  }
  return k;
}

template <class K>
__global__ __launch_bounds__(kSortThreads) void k_count(const K *__restrict__ keys, int64_t n,
                                                         int shift, int ntiles,
                                                         unsigned *__restrict__ counts,
                                                         const unsigned *__restrict__ skip,
                                                         const SegTile *__restrict__ tiles,
                                                         const SceneParams *__restrict__ sp =
                                                             nullptr) {
  if (skip && skip[shift >> 3]) return;
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int64_t base;
  int cnt;
  tile_of(tiles, n, base, cnt);
  const KeySource src = key_source(tiles, sp, n);
  keys += src.offset;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int j = r * kSortThreads + threadIdx.x;
    if (j < cnt)
      atomicAdd(&h[(unsigned)(source_key(keys[base + j], src.p) >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[count_index(tiles, ntiles, threadIdx.x)] = h[threadIdx.x];
}

// ---- stable scatter -----------------------------------------------------------
template <class K>
__global__ __launch_bounds__(kSortThreads) void k_scatter(const K *__restrict__ kin,
                                                           const uint32_t *__restrict__ vin,
                                                           K *__restrict__ kout,
                                                           uint32_t *__restrict__ vout, int64_t n,
                                                           int shift, int ntiles,
                                                           const unsigned *__restrict__ offsets,
                                                           const unsigned *__restrict__ skip,
                                                           const SegTile *__restrict__ tiles,
                                                           const SceneParams *__restrict__ sp =
                                                               nullptr) {
  int64_t tile0;
  int tcnt;
  tile_of(tiles, n, tile0, tcnt);
  const KeySource src = key_source(tiles, sp, n);
  kin += src.offset;  // vin null: the values are positions
  if (skip && skip[shift >> 3]) {
    // the digit is equal for every key: the stable order is the identity
    for (int j = threadIdx.x; j < tcnt; j += kSortThreads) {
      kout[tile0 + j] = source_key(kin[tile0 + j], src.p);
      vout[tile0 + j] = vin ? vin[tile0 + j] : (uint32_t)(tile0 + j);
    }
    return;
  }
  __shared__ K sk[kTile];
  __shared__ uint32_t sv[kTile];
  __shared__ unsigned wh[kWavesPerBlock][256];  // per-wave running digit counts
  __shared__ unsigned tstart[256];              // digit start inside the tile
  __shared__ unsigned gbase[256];               // global destination of the digit run

  const int t = threadIdx.x;
  const int w = t / kWave;
  const int lane = t % kWave;
  for (int i = t; i < kWavesPerBlock * 256; i += kSortThreads) (&wh[0][0])[i] = 0;
  __syncthreads();

  // Wave w ranks tile positions [w*1024, (w+1)*1024) in 16 rounds of 64.
  const int wbase = w * (kTile / kWavesPerBlock);
  K key[kSortItems];
  uint32_t val[kSortItems];
  unsigned rank[kSortItems];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int j = wbase + r * kWave + lane;
    const bool valid = j < tcnt;
    const int64_t i = tile0 + j;
    key[r] = valid ? source_key(kin[i], src.p) : (K)0;
    val[r] = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    const unsigned dig = (unsigned)(key[r] >> shift) & 255u;
    // lanes holding the same digit: intersect the 8 bit-planes
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t plane = __ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? plane : ~plane;
    }
    const unsigned before = wh[w][dig];
    rank[r] = before + (unsigned)__popcll(peers & lt);
    __builtin_amdgcn_wave_barrier();
    const bool leader = valid && ((peers & lt) == 0);
    if (leader) wh[w][dig] = before + (unsigned)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (!valid) rank[r] = 0xffffffffu;
  }
  __syncthreads();
  // digit t: counts per wave -> per-wave prefix; tile digit starts by a block scan
  {
    unsigned c0 = wh[0][t], c1 = wh[1][t], c2 = wh[2][t], c3 = wh[3][t];
    unsigned tot = c0 + c1 + c2 + c3;
    wh[0][t] = 0;
    wh[1][t] = c0;
    wh[2][t] = c0 + c1;
    wh[3][t] = c0 + c1 + c2;
    // exclusive scan of tot over the 256 digits (4 waves x 64 lanes)
    unsigned x = tot;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      unsigned y = __shfl_up(x, o, kWave);
      if (lane >= o) x += y;
    }
    __shared__ unsigned wsum[kWavesPerBlock];
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    unsigned add = 0;
    for (int k = 0; k < w; ++k) add += wsum[k];
    tstart[t] = x - tot + add;
    gbase[t] = offsets[count_index(tiles, ntiles, t)];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (rank[r] != 0xffffffffu) {
      const unsigned dig = (unsigned)(key[r] >> shift) & 255u;
      const unsigned pos = tstart[dig] + wh[w][dig] + rank[r];
      sk[pos] = key[r];
      sv[pos] = val[r];
    }
  }
  __syncthreads();
  for (int j = t; j < tcnt; j += kSortThreads) {
    const K k = sk[j];
    const unsigned dig = (unsigned)(k >> shift) & 255u;
    const unsigned dst = gbase[dig] + (unsigned)j - tstart[dig];
    kout[dst] = k;
    vout[dst] = sv[j];
  }
}

// ---- exclusive scan ---------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class T>
__device__ inline T block_exclusive_scan(T v, T *total) {
  __shared__ T wsum[kScanThreads / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  T x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) wsum[w] = x;
  __syncthreads();
  T add = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / kWave; ++k) {
    if (k < w) add += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return x - v + add;
}

// n_dev (optional): the device-side count; n is then the host's upper bound.
__device__ inline int64_t scan_len(int64_t n, const int64_t *n_dev) {
  if (!n_dev) return n;
  const int64_t m = *n_dev;
  return m < n ? (m > 0 ? m : 0) : n;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                               int64_t n,
                                                               const int64_t *__restrict__ n_dev,
                                                               T *__restrict__ partial) {
  n = scan_len(n, n_dev);
  if ((int64_t)blockIdx.x * kScanTile >= n) {  // beyond the count: never read
    if (threadIdx.x == 0) partial[blockIdx.x] = 0;
    return;
  }
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += (T)in[base + k];
  T tot;
  block_exclusive_scan<T>(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(T *__restrict__ partial, int64_t m,
                                                                 int64_t n,
                                                                 const int64_t *__restrict__ n_dev) {
  // m block partials; only those of blocks below the count are scanned, and
  // the total goes to partial[m]
  const int64_t live = (scan_len(n, n_dev) + kScanTile - 1) / kScanTile;
  const int64_t ml = live < m ? live : m;
  T carry = 0;
  for (int64_t base = 0; base < ml; base += kScanThreads) {
    int64_t i = base + threadIdx.x;
    T v = i < ml ? partial[i] : (T)0;
    T tot;
    T ex = block_exclusive_scan<T>(v, &tot);
    if (i < ml) partial[i] = ex + carry;
    carry += tot;
  }
  if (threadIdx.x == 0) partial[m] = carry;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t *__restrict__ in, int64_t n,
                                                             const int64_t *__restrict__ n_dev,
                                                             const T *__restrict__ partial,
                                                             T *__restrict__ out) {
  const int64_t nb = gridDim.x;
  const int64_t cap = n;
  n = scan_len(n, n_dev);
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // the total, at the count and at the bound
    out[n] = partial[nb];
    if (n != cap) out[cap] = partial[nb];
  }
  if ((int64_t)blockIdx.x * kScanTile >= n) return;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? (T)in[base + k] : (T)0;
    s += v[k];
  }
  T tot;
  T ex = block_exclusive_scan<T>(s, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

template <class T>
T *scan_scratch_reserve(int64_t n, DevBuf &tmp) {
  const int64_t nb = n > 0 ? (n + kScanTile - 1) / kScanTile : 1;
  return tmp.as<T>((size_t)nb + 1);
}

template <class T>
void scan_impl(const uint32_t *in, T *out, int64_t n, DevBuf &tmp, hipStream_t st,
               const int64_t *n_dev = nullptr) {
  const int64_t nb = n > 0 ? (n + kScanTile - 1) / kScanTile : 1;
  T *partial = scan_scratch_reserve<T>(n, tmp);
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n, n_dev,
                     partial);
  hipLaunchKernelGGL(k_scan_partials<T>, dim3(1), dim3(kScanThreads), 0, st, partial, nb, n, n_dev);
  hipLaunchKernelGGL(k_scan_down<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n, n_dev,
                     (const T *)partial, out);
  STQ_HIP(hipGetLastError());
}

// Digit d is skipped when one bucket of its histogram holds all n keys.
__global__ __launch_bounds__(kSortThreads) void k_skip_flags(const unsigned *__restrict__ hist,
                                                              int ndigits, int64_t n,
                                                              unsigned *__restrict__ skip) {
  if (threadIdx.x < 8) skip[threadIdx.x] = 0;
  __syncthreads();
  for (int d = 0; d < ndigits; ++d)
    if ((int64_t)hist[d * 256 + threadIdx.x] == n) skip[d] = 1;
}

__global__ __launch_bounds__(kSortThreads) void k_zero_u32(unsigned *__restrict__ p, int n) {
  const int i = blockIdx.x * kSortThreads + threadIdx.x;
  if (i < n) p[i] = 0;
}

// With sync_free, the skip decision is taken on the device (a skipped digit
// costs one copy pass) and nothing is read back to the host, so the sequence
// can be captured in a graph; otherwise the histograms are read back and
// skipped digits cost nothing.
template <class K>
void radix_sort_impl(K *keys, uint32_t *vals, int64_t n, int bits, SortScratch &s,
                     hipStream_t st, int *passes_run, bool sync_free, bool all_digits = false) {
  if (passes_run) *passes_run = 0;
  if (n <= 1) return;
  const int ndigits = (bits + 7) / 8;
  unsigned *hist = s.hist.as<unsigned>(8 * 256 + 8);
  unsigned *skip = hist + 8 * 256;
  const int64_t ntiles = (n + kTile - 1) / kTile;
  K *ak = s.alt_keys.as<K>((size_t)n);
  uint32_t *av = s.alt_vals.as<uint32_t>((size_t)n);
  unsigned *counts = s.counts.as<unsigned>((size_t)(256 * ntiles + 1));
  scan_scratch_reserve<unsigned>(256 * ntiles, s.misc);
  unsigned hh[8 * 256];
  // Digit histograms find the digits that are equal for all keys: skipped
  // (read back to the host) or made copy passes (flags on the device).  With
  // all_digits there is no histogram and every digit is a pass.
  if (!all_digits) {
    hipLaunchKernelGGL(k_zero_u32, dim3(8), dim3(kSortThreads), 0, st, hist, 8 * 256);
    const int hb = (int)std::min<int64_t>((n + kSortThreads - 1) / kSortThreads, 2048);
    hipLaunchKernelGGL(k_hist_all<K>, dim3(hb), dim3(kSortThreads), 0, st, (const K *)keys, n,
                       ndigits, hist);
    STQ_HIP(hipGetLastError());
    if (sync_free) {
      hipLaunchKernelGGL(k_skip_flags, dim3(1), dim3(kSortThreads), 0, st,
                         (const unsigned *)hist, ndigits, n, skip);
    } else {
      STQ_HIP(hipMemcpyAsync(hh, hist, sizeof(unsigned) * ndigits * 256, hipMemcpyDeviceToHost,
                             st));
      STQ_HIP(hipStreamSynchronize(st));
    }
  }

  K *src_k = keys, *dst_k = ak;
  uint32_t *src_v = vals, *dst_v = av;
  int run = 0;
  for (int d = 0; d < ndigits; ++d) {
    if (!sync_free && !all_digits) {
      bool trivial = false;
      for (int b = 0; b < 256; ++b)
        if ((int64_t)hh[d * 256 + b] == n) trivial = true;
      if (trivial) continue;
    }
    const int shift = 8 * d;
    const unsigned *sk = sync_free && !all_digits ? skip : nullptr;
    hipLaunchKernelGGL(k_count<K>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, st,
                       (const K *)src_k, n, shift, (int)ntiles, counts, sk,
                       (const SegTile *)nullptr);
    scan_impl<unsigned>(counts, counts, 256 * ntiles, s.misc, st);
    hipLaunchKernelGGL(k_scatter<K>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, st,
                       (const K *)src_k, (const uint32_t *)src_v, dst_k, dst_v, n, shift,
                       (int)ntiles, (const unsigned *)counts, sk, (const SegTile *)nullptr);
    STQ_HIP(hipGetLastError());
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
    ++run;
  }
  if (src_k != keys) {
    STQ_HIP(hipMemcpyAsync(keys, src_k, sizeof(K) * n, hipMemcpyDeviceToDevice, st));
    STQ_HIP(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
  }
  if (passes_run) *passes_run = run;
}

}  // namespace

// Every digit a pass, tiles from the segment table: the keys of each segment
// are sorted stably among themselves and stay in the segment.
void radix_sort_u32_segmented(uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                              const SegTile *tiles, int64_t ntiles, SortScratch &s,
                              hipStream_t st) {
  if (n <= 1 || ntiles <= 0) return;
  const int ndigits = (bits + 7) / 8;
  uint32_t *ak = s.alt_keys.as<uint32_t>((size_t)n);
  uint32_t *av = s.alt_vals.as<uint32_t>((size_t)n);
  unsigned *counts = s.counts.as<unsigned>((size_t)(256 * ntiles + 1));
  scan_scratch_reserve<unsigned>(256 * ntiles, s.misc);
  uint32_t *src_k = keys, *dst_k = ak, *src_v = vals, *dst_v = av;
  for (int d = 0; d < ndigits; ++d) {
    const int shift = 8 * d;
    hipLaunchKernelGGL(k_count<uint32_t>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, st,
                       (const uint32_t *)src_k, n, shift, (int)ntiles, counts,
                       (const unsigned *)nullptr, tiles);
    scan_impl<unsigned>(counts, counts, 256 * ntiles, s.misc, st);
    hipLaunchKernelGGL(k_scatter<uint32_t>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, st,
                       (const uint32_t *)src_k, (const uint32_t *)src_v, dst_k, dst_v, n, shift,
                       (int)ntiles, (const unsigned *)counts, (const unsigned *)nullptr, tiles);
    STQ_HIP(hipGetLastError());
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
  }
  if (src_k != keys) {
    STQ_HIP(hipMemcpyAsync(keys, src_k, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    STQ_HIP(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
  }
}
void radix_sort_u32_segmented_axes(uint32_t *keys3, uint32_t *vals, int64_t n, int bits,
                                   const SegTile *tiles, int64_t ntiles, const SceneParams *sp,
                                   SortScratch &s, hipStream_t st) {
  if (n <= 0 || ntiles <= 0) return;
  const int ndigits = (bits + 7) / 8;
  uint32_t *ak = s.alt_keys.as<uint32_t>((size_t)n);
  uint32_t *av = s.alt_vals.as<uint32_t>((size_t)n);
  unsigned *counts = s.counts.as<unsigned>((size_t)(256 * ntiles + 1));
  scan_scratch_reserve<unsigned>(256 * ntiles, s.misc);
  uint32_t *src_k = keys3, *dst_k = ak, *src_v = nullptr, *dst_v = av;
  for (int d = 0; d < ndigits; ++d) {
    const int shift = 8 * d;
    const SceneParams *ps = d == 0 ? sp : nullptr;
    hipLaunchKernelGGL(k_count<uint32_t>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, st,
                       (const uint32_t *)src_k, n, shift, (int)ntiles, counts,
                       (const unsigned *)nullptr, tiles, ps);
    scan_impl<unsigned>(counts, counts, 256 * ntiles, s.misc, st);
    hipLaunchKernelGGL(k_scatter<uint32_t>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, st,
                       (const uint32_t *)src_k, (const uint32_t *)src_v, dst_k, dst_v, n, shift,
                       (int)ntiles, (const unsigned *)counts, (const unsigned *)nullptr, tiles, ps);
    STQ_HIP(hipGetLastError());
    if (d == 0) {  // from here on the keys ping-pong between alt and keys3[0 .. n)
      src_k = ak;
      dst_k = keys3;
      src_v = av;
      dst_v = vals;
    } else {
      std::swap(src_k, dst_k);
      std::swap(src_v, dst_v);
    }
  }
  if (src_k != keys3) {
    STQ_HIP(hipMemcpyAsync(keys3, src_k, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    STQ_HIP(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
  }
}

void radix_sort_segmented_reserve(int64_t n, int64_t ntiles, SortScratch &s) {
  if (n <= 1) return;
  s.alt_keys.as<uint32_t>((size_t)n);
  s.alt_vals.as<uint32_t>((size_t)n);
  s.counts.as<unsigned>((size_t)(256 * ntiles + 1));
  scan_scratch_reserve<unsigned>(256 * ntiles, s.misc);
}
int sort_tile_keys() { return kTile; }

void radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bits, SortScratch &s,
                    hipStream_t st, int *passes_run) {
  radix_sort_impl<uint64_t>(keys, vals, n, bits, s, st, passes_run, false);
}
void radix_sort_u32(uint32_t *keys, uint32_t *vals, int64_t n, int bits, SortScratch &s,
                    hipStream_t st, int *passes_run) {
  radix_sort_impl<uint32_t>(keys, vals, n, bits, s, st, passes_run, false);
}
void radix_sort_u64_sync_free(uint64_t *keys, uint32_t *vals, int64_t n, int bits,
                              SortScratch &s, hipStream_t st) {
  radix_sort_impl<uint64_t>(keys, vals, n, bits, s, st, nullptr, true);
}
void radix_sort_u32_sync_free(uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                              SortScratch &s, hipStream_t st) {
  radix_sort_impl<uint32_t>(keys, vals, n, bits, s, st, nullptr, true);
}
void radix_sort_u32_all_digits(uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                               SortScratch &s, hipStream_t st) {
  radix_sort_impl<uint32_t>(keys, vals, n, bits, s, st, nullptr, true, true);
}
void radix_sort_reserve(int64_t n, int key_bytes, SortScratch &s) {
  if (n <= 1) return;
  const int64_t ntiles = (n + kTile - 1) / kTile;
  s.hist.as<unsigned>(8 * 256 + 8);
  s.alt_keys.get((size_t)n * key_bytes);
  s.alt_vals.as<uint32_t>((size_t)n);
  s.counts.as<unsigned>((size_t)(256 * ntiles + 1));
  scan_scratch_reserve<unsigned>(256 * ntiles, s.misc);
}
void scan_reserve_i64(int64_t n, DevBuf &tmp) { scan_scratch_reserve<int64_t>(n, tmp); }
void scan_u32_to_i64(const uint32_t *in, int64_t *out, int64_t n, DevBuf &tmp, hipStream_t st,
                     const int64_t *n_dev) {
  scan_impl<int64_t>(in, out, n, tmp, st, n_dev);
}
void scan_u32(const uint32_t *in, uint32_t *out, int64_t n, DevBuf &tmp, hipStream_t st) {
  scan_impl<unsigned>(in, out, n, tmp, st);
}

int device_cu_count(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    (void)hipGetLastError();
    return 256;
  }
  return v > 0 ? v : 256;
}

}  // namespace stqamd
