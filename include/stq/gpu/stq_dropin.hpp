// stq_dropin.hpp -- C++ drop-in for the reference GPU broad-phase entry.
//
// Replaces, for host code built without CUDA:
//   #include <stq/gpu/simulation.cuh>      (reference src/simulation.cuh:22-27)
//   stq::gpu::Aabb                          (reference src/aabb.cuh:15-43, FP64)
//   stq::gpu::MemHandler                    (reference src/memory.cuh:69-191)
//   stq::gpu::run_sweep_sharedqueue(...)    (reference src/simulation.cu:555-782)
// with the same names, argument meaning and output convention, implemented
// over the C-ABI of <stq.h> (libstq_amd.so, gfx950).  No CUDA or HIP type
// appears here: int2/int3 are plain structs of this namespace.
//
// As in the reference, one call sweeps the sorted rows [tidstart, tidstart +
// MAX_OVERLAP_CUTOFF) (src/sweep.cu:234-240) and advances tidstart by the
// cutoff (src/simulation.cu:747), so a caller that loops while tidstart < N
// gets every pair exactly once.  The output is sized as the reference sizes
// it (src/memory.cuh:78-94,163-188): room for at most
//   (allocatable - sizeof(CCDConfig)) / (sizeof(CCDData) + 3 * sizeof(int2))
// pairs, allocatable = min(0.95 * free device memory, limitGB GiB - memory in
// use) with limitGB = memlimit when given; while the range holds more pairs
// than that, the cutoff is halved and the range counted again.
// Differences a caller may notice (all in the direction of the reference's
// documented intent):
//  * limitGB = 0 means no limit, and a limit below the memory in use leaves
//    no room (the reference's size_t subtraction wraps around there);
//  * a cutoff is never halved below one row; a single row with more pairs
//    than fit is an error;
//  * d_overlaps holds exactly *d_count pairs, each (smaller id, larger id),
//    ids being input positions as in the reference (src/sweep.cu:160);
//  * the pair set is exactly the FP64 set (the reference's default build
//    rounds boxes to float outward and returns a superset).
// On failure it prints the message and calls std::exit with the status, as
// the reference's gpuErrchk does (src/memory.cuh:11-19).
#pragma once

#include <stq.h>

#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

namespace stq::gpu {

struct int2 {
  int x, y;
};
struct int3 {
  int x, y, z;
};
struct double3 {
  double x, y, z;
};

// The reference's FP64 box (Scalar = double), byte for byte.
class Aabb {
public:
  int id;
  double3 min;
  double3 max;
  int3 vertexIds;
  int ref_id;

  Aabb() = default;
  Aabb(int assignid, int reference_id, const int *vids, const double *tempmin,
       const double *tempmax)
      : id(assignid), min{tempmin[0], tempmin[1], tempmin[2]},
        max{tempmax[0], tempmax[1], tempmax[2]}, vertexIds{vids[0], vids[1], vids[2]},
        ref_id(reference_id) {}
};
static_assert(sizeof(Aabb) == sizeof(stq_aabb_f64), "Aabb must match the FP64 reference layout");

// The reference's broad-phase sizing state (src/memory.cuh:69-76).
struct MemHandler {
  size_t MAX_OVERLAP_CUTOFF = 0;
  size_t MAX_OVERLAP_SIZE = 0;
  size_t MAX_UNIT_SIZE = 0;
  size_t MAX_QUERIES = 0;
  int realcount = 0;
  int limitGB = 0;
};

// Sizes of the reference's narrow-phase records that its output sizing
// reserves room for (src/memory.cuh:21-67): CCDData is the 256-byte query
// record (stq_ccd_query); CCDConfig is 64 bytes on its build.
constexpr size_t kCCDDataBytes = 256;
constexpr size_t kCCDConfigBytes = 64;

namespace detail {
inline void *context() {
  static void *ctx = nullptr;
  if (!ctx) {
    stq_options opt;
    stq_options_default(&opt);
    const int rc = stq_create(&ctx, 0, &opt);
    if (rc != STQ_OK) {
      std::fprintf(stderr, "stq_create failed: %s\n", stq_status_string(rc));
      std::exit(rc);
    }
  }
  return ctx;
}
inline void check(void *ctx, int rc, const char *what) {
  if (rc != STQ_OK) {
    std::fprintf(stderr, "%s failed: %s: %s\n", what, stq_status_string(rc),
                 stq_last_error(ctx));
    std::exit(rc);
  }
}
}  // namespace detail

// Device memory returned by run_sweep_sharedqueue; the reference's caller
// releases it with cudaFree.
inline void free_device(void *d_ptr) { stq_free_device(detail::context(), d_ptr); }

namespace detail {
// The reference's allocatable bytes (MemHandler::__getAllocatable) without its
// wrap-around: limitGB 0 is no limit.
inline size_t allocatable(void *ctx, const MemHandler &mh) {
  size_t free_b = 0, total_b = 0;
  check(ctx, stq_device_memory(ctx, &free_b, &total_b), "stq_device_memory");
  const size_t used = total_b - free_b;
  const size_t def = (size_t)(0.95 * (double)free_b);
  if (mh.limitGB <= 0) return def;
  const size_t lim = (size_t)mh.limitGB << 30;
  const size_t user = lim > used ? lim - used : 0;
  return user < def ? user : def;
}
// MemHandler::__getLargestOverlap: pairs that fit with their narrow-phase room.
inline size_t largest_overlap(void *ctx, const MemHandler &mh) {
  const size_t a = allocatable(ctx, mh);
  return a > kCCDConfigBytes ? (a - kCCDConfigBytes) / (kCCDDataBytes + 3 * sizeof(int2)) : 0;
}
// Pairs of the sorted rows [r0, r1), without writing them.
inline int64_t count_rows(void *ctx, const void *d_boxes, int N, int64_t r0, int64_t r1) {
  int64_t count = 0;
  const int rc = stq_find_pairs_rows_device(ctx, static_cast<const stq_aabb_f64 *>(d_boxes), N, r0,
                                            r1, nullptr, 0, &count);
  if (rc != STQ_ERR_CAPACITY) check(ctx, rc, "stq_find_pairs_rows_device");
  return count;
}
}  // namespace detail

// Same signature and meaning as the reference (src/simulation.cuh:22-27):
// boxes: host array of N boxes (copied in, src/simulation.cu:578-580);
// memhandle: MAX_OVERLAP_CUTOFF (0: N) bounds the rows of this call and may be
// halved, MAX_OVERLAP_SIZE and realcount are set, limitGB = memlimit when
// memlimit != 0 (src/simulation.cu:563-568); d_overlaps / d_count: allocated
// here and owned by the caller (release with free_device), d_count holding
// the pair count as the reference's realcount (src/simulation.cu:752-753);
// threads: set to the sweep's block size; tidstart: the first sorted row of
// this call, advanced by the cutoff; nbox, devcount: accepted and unused.
// With KEEP_CPU_OVERLAPS defined the pairs are also appended to finOverlaps
// (src/simulation.cu:763-780).
inline void run_sweep_sharedqueue(const Aabb *boxes, MemHandler *memhandle, int N, int nbox,
                                  std::vector<std::pair<int, int>> &finOverlaps,
                                  int2 *&d_overlaps, int *&d_count, int &threads,
                                  int &tidstart, int &devcount, const int memlimit = 0) {
  (void)nbox;
  (void)devcount;
  (void)finOverlaps;
  void *ctx = detail::context();
  MemHandler local;
  MemHandler &mh = memhandle ? *memhandle : local;
  if (!mh.MAX_OVERLAP_CUTOFF) mh.MAX_OVERLAP_CUTOFF = (size_t)(N > 0 ? N : 1);
  if (memlimit) mh.limitGB = memlimit;
  void *d_boxes = nullptr;
  detail::check(ctx, stq_device_alloc(ctx, sizeof(Aabb) * (size_t)(N > 0 ? N : 1), &d_boxes),
                "stq_device_alloc");
  detail::check(ctx, stq_copy_to_device(ctx, d_boxes, boxes, sizeof(Aabb) * (size_t)N),
                "stq_copy_to_device");
  const int64_t r0 = tidstart < 0 ? 0 : tidstart;
  auto range_end = [&]() { return r0 + (int64_t)mh.MAX_OVERLAP_CUTOFF; };
  int64_t count = detail::count_rows(ctx, d_boxes, N, r0, range_end());
  mh.MAX_OVERLAP_SIZE = detail::largest_overlap(ctx, mh);  // setOverlapSize
  while ((size_t)count > mh.MAX_OVERLAP_SIZE) {             // handleBroadPhaseOverflow
    const size_t largest = detail::largest_overlap(ctx, mh);
    mh.MAX_OVERLAP_SIZE = largest < (size_t)count ? largest : (size_t)count;
    if (mh.MAX_OVERLAP_SIZE >= (size_t)count) break;
    if (mh.MAX_OVERLAP_CUTOFF <= 1) {
      std::fprintf(stderr, "run_sweep_sharedqueue: the pairs of one row do not fit in memory\n");
      std::exit(STQ_ERR_OUT_OF_MEMORY);
    }
    mh.MAX_OVERLAP_CUTOFF /= 2;
    count = detail::count_rows(ctx, d_boxes, N, r0, range_end());
  }
  void *dp = nullptr;
  detail::check(ctx, stq_device_alloc(ctx, sizeof(stq_pair) * (size_t)(count > 0 ? count : 1), &dp),
                "stq_device_alloc");
  int64_t written = 0;
  detail::check(ctx, stq_find_pairs_rows_device(ctx, static_cast<const stq_aabb_f64 *>(d_boxes), N, r0,
                                                range_end(), static_cast<stq_pair *>(dp), count,
                                                &written),
                "stq_find_pairs_rows_device");
  stq_free_device(ctx, d_boxes);
  const int c32 = (int)written;
  void *dc = nullptr;
  detail::check(ctx, stq_device_alloc(ctx, sizeof(int), &dc), "stq_device_alloc");
  detail::check(ctx, stq_copy_to_device(ctx, dc, &c32, sizeof(int)), "stq_copy_to_device");
  d_overlaps = static_cast<int2 *>(dp);
  d_count = static_cast<int *>(dc);
#ifdef KEEP_CPU_OVERLAPS
  std::vector<stq_pair> host((size_t)written);
  detail::check(ctx, stq_copy_to_host(ctx, host.data(), dp, sizeof(stq_pair) * (size_t)written),
                "stq_copy_to_host");
  finOverlaps.reserve(finOverlaps.size() + (size_t)written);
  for (const stq_pair &p : host) finOverlaps.emplace_back(p.first, p.second);
#endif
  mh.realcount = c32;
  threads = 256;
  tidstart = (int)(r0 + (int64_t)mh.MAX_OVERLAP_CUTOFF);
}

}  // namespace stq::gpu
