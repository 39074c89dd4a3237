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
// Differences a caller may notice (all in the direction of the reference's
// documented intent):
//  * the whole scene is processed in one call: MAX_OVERLAP_CUTOFF is set to N
//    and tidstart advances by N, so a caller's loop over tidstart ends after
//    one call (the reference can leave boxes unprocessed when its cutoff
//    shrinks, src/memory.cuh:183-187);
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

// The reference's broad-phase sizing state.  Only the fields remain: the
// library sizes its output exactly, so no memory probing is needed.
struct MemHandler {
  size_t MAX_OVERLAP_CUTOFF = 0;
  size_t MAX_OVERLAP_SIZE = 0;
  size_t MAX_UNIT_SIZE = 0;
  size_t MAX_QUERIES = 0;
  int realcount = 0;
  int limitGB = 0;
};

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

// Same signature and meaning as the reference (src/simulation.cuh:22-27):
// boxes: host array of N boxes (copied in, src/simulation.cu:578-580);
// d_overlaps / d_count: allocated here and owned by the caller (release with
// free_device), d_count holding the pair count as the reference's realcount
// (src/simulation.cu:752-753); threads: set to the sweep's block size;
// tidstart: advanced past the processed boxes; nbox, devcount, memlimit:
// accepted and unused.  With KEEP_CPU_OVERLAPS defined the pairs are also
// appended to finOverlaps (src/simulation.cu:763-780).
inline void run_sweep_sharedqueue(const Aabb *boxes, MemHandler *memhandle, int N, int nbox,
                                  std::vector<std::pair<int, int>> &finOverlaps,
                                  int2 *&d_overlaps, int *&d_count, int &threads,
                                  int &tidstart, int &devcount, const int memlimit = 0) {
  (void)nbox;
  (void)devcount;
  (void)memlimit;
  (void)finOverlaps;
  void *ctx = detail::context();
  if (memhandle && !memhandle->MAX_OVERLAP_CUTOFF) memhandle->MAX_OVERLAP_CUTOFF = (size_t)N;
  // boxes before tidstart were handled by an earlier call
  const int n = tidstart < N ? N : 0;
  void *d_boxes = nullptr;
  detail::check(ctx, stq_device_alloc(ctx, sizeof(Aabb) * (size_t)(n > 0 ? n : 1), &d_boxes),
                "stq_device_alloc");
  detail::check(ctx, stq_copy_to_device(ctx, d_boxes, boxes, sizeof(Aabb) * (size_t)n),
                "stq_copy_to_device");
  stq_pair *d_pairs = nullptr;
  int64_t count = 0;
  detail::check(ctx, stq_find_pairs_device_alloc(ctx, static_cast<const stq_aabb_f64 *>(d_boxes), n,
                                                 &d_pairs, &count),
                "stq_find_pairs_device_alloc");
  stq_free_device(ctx, d_boxes);
  const int c32 = (int)count;
  void *dc = nullptr;
  detail::check(ctx, stq_device_alloc(ctx, sizeof(int), &dc), "stq_device_alloc");
  detail::check(ctx, stq_copy_to_device(ctx, dc, &c32, sizeof(int)), "stq_copy_to_device");
  d_overlaps = reinterpret_cast<int2 *>(d_pairs);
  d_count = static_cast<int *>(dc);
#ifdef KEEP_CPU_OVERLAPS
  std::vector<stq_pair> host((size_t)count);
  detail::check(ctx, stq_copy_to_host(ctx, host.data(), d_pairs, sizeof(stq_pair) * (size_t)count),
                "stq_copy_to_host");
  finOverlaps.reserve(finOverlaps.size() + (size_t)count);
  for (const stq_pair &p : host) finOverlaps.emplace_back(p.first, p.second);
#endif
  if (memhandle) {
    memhandle->MAX_OVERLAP_SIZE = (size_t)count;
    memhandle->realcount = c32;
  }
  threads = 256;
  tidstart = N;
}

}  // namespace stq::gpu
