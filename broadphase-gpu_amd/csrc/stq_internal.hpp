// stq_internal.hpp -- shared definitions of the gfx950 STQ broad phase.
//
// Device-side data layout (all by sorted position s along the sweep axis):
//   orig[s]    int32   input position of the box at s (the sort's values)
//   smin[s]    T       sweep-axis minimum (T = double, or float rounded down)
//   smax[s]    T       sweep-axis maximum (T = double, or float rounded up)
//   geo[s]     4 x T   {min, max} on the two other axes
//   meta[s]    int4    {v0, v1, v2, typebit}; typebit 1 vertex, 2 edge, 4 face
//   full[s]    6 x f64 the FP64 box (FP32-confirm form only)
//   hi[s]      int32   last sorted position q with smin[q] <= smax[s]
// The reference keeps the same split as Scalar2 sm[] plus MiniBox mini[]
// (src/sweep.cu:125-161); here the other-axis bounds and the ids are separate
// arrays so that the sweep's hot loop fetches 16 or 32 bytes per step.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/stq.h"

namespace stqamd {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 XCDs of 32 CUs

// Most blocks a grid-stride kernel that ends in one atomic per block launches
// (8 blocks of 256 threads per CU on 256 CUs): atomics to one address
// serialise in one L2 channel, so a counter takes at most this many.
constexpr unsigned kReduceGrid = 2048;

// Adds the sum of v over the block (NT threads, all of them call this) to
// *dst with one atomic: wave sums by shuffles, then the block's waves in LDS.
template <int NT, class T>
__device__ __forceinline__ void block_atomic_add(T v, T *dst) {
  __shared__ T part[NT / kWave];
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = part[0];
    for (int w = 1; w < NT / kWave; ++w) s += part[w];
    if (s != T(0)) atomicAdd(dst, s);
  }
}

// Type bits of a vertex-id triple (src/aabb.cu:12-30).
__host__ __device__ inline int type_bit(int v1, int v2) {
  return v2 >= 0 ? 4 : (v1 >= 0 ? 2 : 1);
}
// Type bits a row of type `bit` pairs with (is_valid_pair, src/aabb.cu:37-40):
// vertex-face, face-vertex, edge-edge.
__host__ __device__ inline int accept_bits(int bit) {
  return bit == 1 ? 4 : (bit == 4 ? 1 : 2);
}

// Order-preserving unsigned images of IEEE values (negative values have all
// bits flipped, non-negative values the sign bit set).  -0.0 is mapped like
// +0.0 so that equal values have equal keys.
__host__ __device__ inline uint64_t key_of(double d) {
  d = d + 0.0;
  uint64_t u;
  __builtin_memcpy(&u, &d, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline uint32_t key_of(float f) {
  f = f + 0.0f;
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// FP32 outward rounding (src/aabb.cu:56-61): round to nearest, then one ulp
// toward -inf / +inf, i.e. nextafterf(f, -inf) / nextafterf(f, +inf) written
// on the bit pattern.  Inputs are never NaN (validated before).
__host__ __device__ inline float next_down(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) == 0) u = 0x80000001u;          // +-0 -> -denorm_min
  else if (u == 0xff800000u) {}                          // -inf stays
  else if (u & 0x80000000u) u += 1;                      // negative: away from 0
  else u -= 1;                                           // positive (incl. +inf)
  float r;
  __builtin_memcpy(&r, &u, 4);
  return r;
}
__host__ __device__ inline float next_up(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) == 0) u = 0x00000001u;          // +-0 -> +denorm_min
  else if (u == 0x7f800000u) {}                          // +inf stays
  else if (u & 0x80000000u) u -= 1;                      // negative (incl. -inf)
  else u += 1;
  float r;
  __builtin_memcpy(&r, &u, 4);
  return r;
}
__host__ __device__ inline float round_down_f32(double d) { return next_down((float)d); }
__host__ __device__ inline float round_up_f32(double d) { return next_up((float)d); }

struct Item {   // one sweep work item: up to 64 rows against a range of q
  int rowoff;   // first entry of the group's rows in rowlist[] ({position, chain end})
  int nrows;    // rows in the group (<= 64)
  int q0, q1;   // inclusive q range of this item
};

// Error flags raised by kernels.
enum : unsigned { kErrNaN = 1u, kErrInverted = 2u };

#define STQ_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) throw ::stqamd::HipError(e_, #call, __FILE__, __LINE__); \
  } while (0)

struct HipError {
  hipError_t code;
  std::string what;
  HipError(hipError_t c, const char *call, const char *file, int line) : code(c) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", call, hipGetErrorString(c), file, line);
    what = buf;
  }
};

struct StatusError {
  int status;
  std::string what;
  StatusError(int s, std::string w) : status(s), what(std::move(w)) {}
};

// Incremented whenever any DevBuf is (re)allocated or released: a captured
// graph that holds buffer addresses is valid only while it is unchanged.
inline std::atomic<uint64_t> g_devbuf_generation{0};

// A fixed device region that buffers are carved from in order (256-byte
// aligned), instead of being allocated: used by the allocation-free entry.
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
};

// Grow-only device buffer.  With an arena, growing carves a new piece of it
// (nothing is allocated; running out is STQ_ERR_TOO_LARGE).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  Arena *arena = nullptr;
  void *get(size_t need) {
    if (need == 0) need = 256;
    if (need > bytes && arena) {
      const size_t at = (arena->used + 255) & ~(size_t)255;
      if (at > arena->size || need > arena->size - at)
        throw StatusError(STQ_ERR_TOO_LARGE, "the call needs more scratch than the fixed array "
                                             "of the allocation-free entry holds");
      p = arena->base + at;
      bytes = need;
      arena->used = at + need;
      return p;
    }
    if (need > bytes) {
      g_devbuf_generation.fetch_add(1);
      if (p) (void)hipFree(p);
      p = nullptr;
      bytes = 0;
      size_t want = need + need / 8;
      if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        if (hipMalloc(&p, need) != hipSuccess) {
          (void)hipGetLastError();
          p = nullptr;
          throw StatusError(STQ_ERR_OUT_OF_MEMORY, "device allocation failed");
        }
        want = need;
      }
      bytes = want;
    }
    return p;
  }
  template <class T> T *as(size_t count) { return static_cast<T *>(get(count * sizeof(T))); }
  void release() {
    if (p && !arena) {
      g_devbuf_generation.fetch_add(1);
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
};

// ---- radix sort / scan (radix_sort.hip) ---------------------------------
// Stable LSD radix sort of (key, value) pairs by `bits` low key bits;
// digits whose value is the same for every key are skipped.  Result is in
// keys/vals (the alternate buffers are scratch).
struct SortScratch {
  DevBuf alt_keys, alt_vals, counts, hist, misc;
};
void radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bits,
                    SortScratch &s, hipStream_t st, int *passes_run);
void radix_sort_u32(uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                    SortScratch &s, hipStream_t st, int *passes_run);
// The same sort with no host synchronisation (capturable in a graph): a digit
// that is equal for all keys costs a copy pass instead of nothing.
void radix_sort_u64_sync_free(uint64_t *keys, uint32_t *vals, int64_t n, int bits,
                              SortScratch &s, hipStream_t st);
void radix_sort_u32_sync_free(uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                              SortScratch &s, hipStream_t st);
// The sync-free sort without the digit histogram: every digit is a pass, for
// keys whose digits are expected to vary (FP32 images of coordinates).
void radix_sort_u32_all_digits(uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                               SortScratch &s, hipStream_t st);
// Allocates the scratch a sort of n keys of key_bytes needs (nothing is
// allocated by the sort itself afterwards).
void radix_sort_reserve(int64_t n, int key_bytes, SortScratch &s);
// A tile of the segmented sort: keys [start, start + count) of one segment
// (count <= sort_tile_keys()); its count of digit d is counts[cbase + d *
// stride] (cbase = 256 * the segment's first tile + the tile's index in the
// segment, stride = the segment's number of tiles).
struct SegTile {
  int64_t start;
  int64_t cbase;
  int32_t count;
  int32_t stride;
  int32_t scene;  // the segment's scene (picks its sweep axis's keys in the first pass)
  int32_t pad;
};
// Keys per sort tile.
int sort_tile_keys();
// Segmented sort of 4-byte keys (every digit a pass, no host read-back): the
// keys of each segment, whose tiles are listed in `tiles`, are sorted stably
// among themselves and stay in place as a segment.
void radix_sort_u32_segmented(uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                              const SegTile *tiles, int64_t ntiles, SortScratch &s,
                              hipStream_t st);
void radix_sort_segmented_reserve(int64_t n, int64_t ntiles, SortScratch &s);
// The same sort with the FP32 keys of all three axes as input: keys3 holds
// 3 n keys (axis a at keys3[a n ..]); in the first pass tile t reads the keys
// of its scene's sweep axis (sp[scene].axis) and turns each into the scene's
// sweep key (skey_down, sweep.hpp), and the values start as the positions
// 0 .. n-1; `bits` is the width of the sweep keys.  The sorted sweep keys end
// in keys3[0 .. n), their positions in vals.  With tiles null: one scene
// (sp[0]), tiles of sort_tile_keys() keys over all n (ntiles of them).
struct SceneParams;
void radix_sort_u32_segmented_axes(uint32_t *keys3, uint32_t *vals, int64_t n, int bits,
                                   const SegTile *tiles, int64_t ntiles, const SceneParams *sp,
                                   SortScratch &s, hipStream_t st);

// Exclusive prefix sum of n uint32 counts into n+1 int64 offsets (out[n] is
// the total).  Scratch is taken from `tmp`.  With n_dev, only the first
// m = min(*n_dev, n) counts are summed (n is then the host's upper bound):
// out[0..m] are written, the total goes to both out[m] and out[n], and the
// blocks beyond m return at once.
void scan_u32_to_i64(const uint32_t *in, int64_t *out, int64_t n, DevBuf &tmp,
                     hipStream_t st, const int64_t *n_dev = nullptr);
// Allocates the scratch of scan_u32_to_i64 over n counts.
void scan_reserve_i64(int64_t n, DevBuf &tmp);
// Exclusive prefix sum of uint32 into uint32 (n+1 outputs), for sort counts.
void scan_u32(const uint32_t *in, uint32_t *out, int64_t n, DevBuf &tmp,
              hipStream_t st);

int device_cu_count(int device);

}  // namespace stqamd
