// stq_cpu_dropin.hpp -- C++ drop-in for the reference's CPU broad-phase entry.
//
// Replaces, for host code built without TBB, Eigen or spdlog:
//   #include <stq/cpu/sweep.hpp>                (reference src/cpu/sweep.hpp:16-25)
//   stq::cpu::Aabb                              (reference src/cpu/aabb.hpp:21-33, FP64)
//   stq::cpu::sort_along_xaxis, sweep, run_sweep_cpu, sweep_cpu_single_batch
//                                               (reference src/cpu/sweep.cpp:72-158)
// with the same names, argument meaning and results, implemented over the
// C-ABI of <stq.h>: the sort and the sweep run on the GPU (libstq_amd.so,
// gfx950), and this header only moves the caller's boxes and pairs.
//
// Semantics kept from the reference:
//  * sort_along_xaxis sorts the caller's vector by min[0] ascending (ties:
//    earlier vector position first; the reference's tbb::parallel_sort leaves
//    their order unspecified);
//  * sweep(boxes, overlaps, n) reports, for each of the first n boxes a (in
//    vector order, which must be that sorted order), every later box b with
//    a.max[0] >= b.min[0] that overlaps a on the other two axes, forms a
//    vertex-face, face-vertex or edge-edge pair and shares no vertex id, as
//    the pair (a.id, b.id) -- ids are the boxes' id fields;
//  * run_sweep_cpu clears finOverlaps, sweeps the first n boxes and halves n
//    while the pairs do not fit in memory;
//  * sweep_cpu_single_batch sorts when boxes_batching.size() == N, runs
//    run_sweep_cpu, erases the first n boxes and sets n = min(size, n), so
//    the reference's loop of app/main.cpp:147-151 walks the whole scene.
// Differences: the pairs of one call come out in the GPU sweep's fixed order
// (the reference concatenates per-thread lists in thread order); n larger
// than the vector is clamped (the reference reads past its end).
// Failures print a message and call std::exit with the status.
#pragma once

#include <stq.h>

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <utility>
#include <vector>

namespace stq::cpu {

// The reference CPU box built with CCD_USE_DOUBLE.  Its Eigen arrays become
// std::array<double, 3> (min[k] and max[k] read as before); a trailing
// padding word makes the object the 72-byte stq_aabb_f64, so a vector of
// boxes is handed to the library without a copy.
class Aabb {
public:
  int id = 0;
  std::array<double, 3> min{};
  std::array<double, 3> max{};
  std::array<int, 3> vertexIds{};

  Aabb() = default;
  Aabb(int assignid, const std::array<int, 3> &vids, const std::array<double, 3> &tempmin,
       const std::array<double, 3> &tempmax)
      : id(assignid), min(tempmin), max(tempmax), vertexIds(vids) {}

  int pad_ = 0;  // the place of stq_aabb_f64::ref_id; not read
};
static_assert(sizeof(Aabb) == sizeof(stq_aabb_f64), "Aabb must be the 72-byte FP64 box");
static_assert(offsetof(Aabb, min) == offsetof(stq_aabb_f64, min), "min offset");
static_assert(offsetof(Aabb, max) == offsetof(stq_aabb_f64, max), "max offset");
static_assert(offsetof(Aabb, vertexIds) == offsetof(stq_aabb_f64, vertex_ids), "ids offset");

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
    std::fprintf(stderr, "%s failed: %s: %s\n", what, stq_status_string(rc), stq_last_error(ctx));
    std::exit(rc);
  }
}
// A device copy of the caller's boxes, released when it goes out of scope.
struct DeviceBoxes {
  void *ctx;
  void *p = nullptr;
  DeviceBoxes(void *c, const std::vector<Aabb> &boxes) : ctx(c) {
    const size_t bytes = sizeof(Aabb) * (boxes.empty() ? 1 : boxes.size());
    check(ctx, stq_device_alloc(ctx, bytes, &p), "stq_device_alloc");
    check(ctx, stq_copy_to_device(ctx, p, boxes.data(), sizeof(Aabb) * boxes.size()),
          "stq_copy_to_device");
  }
  ~DeviceBoxes() { stq_free_device(ctx, p); }
  const stq_aabb_f64 *get() const { return static_cast<const stq_aabb_f64 *>(p); }
};
}  // namespace detail

// src/cpu/sweep.cpp:72-74: the vector sorted by min[0] (the sweep's own key
// sort on the GPU; the boxes are then permuted here).
inline void sort_along_xaxis(std::vector<Aabb> &boxes) {
  if (boxes.size() < 2) return;
  void *ctx = detail::context();
  const detail::DeviceBoxes d(ctx, boxes);
  void *d_order = nullptr;
  detail::check(ctx, stq_device_alloc(ctx, sizeof(int32_t) * boxes.size(), &d_order),
                "stq_device_alloc");
  std::vector<int32_t> order(boxes.size());
  detail::check(ctx, stq_sort_order_device(ctx, d.get(), (int64_t)boxes.size(),
                                           static_cast<int32_t *>(d_order)),
                "stq_sort_order_device");
  detail::check(ctx, stq_copy_to_host(ctx, order.data(), d_order, sizeof(int32_t) * order.size()),
                "stq_copy_to_host");
  stq_free_device(ctx, d_order);
  std::vector<Aabb> sorted;
  sorted.reserve(boxes.size());
  for (const int32_t i : order) sorted.push_back(boxes[(size_t)i]);
  boxes.swap(sorted);
}

namespace detail {
// The pairs of the first n boxes of a vector in sweep order; false when the
// pairs do not fit in device memory (the caller then halves n).
inline bool sweep_rows(const std::vector<Aabb> &boxes, std::vector<std::pair<int, int>> &overlaps,
                       int n) {
  overlaps.clear();
  if (boxes.empty() || n <= 0) return true;
  void *ctx = context();
  const DeviceBoxes d(ctx, boxes);
  const int64_t N = (int64_t)boxes.size(), rows = std::min<int64_t>(n, N);
  int64_t count = 0;
  int rc = stq_find_pairs_rows_device(ctx, d.get(), N, 0, rows, nullptr, 0, &count);
  if (rc != STQ_ERR_CAPACITY) check(ctx, rc, "stq_find_pairs_rows_device");
  void *d_pairs = nullptr;
  rc = stq_device_alloc(ctx, sizeof(stq_pair) * (size_t)(count > 0 ? count : 1), &d_pairs);
  if (rc == STQ_ERR_OUT_OF_MEMORY) return false;
  check(ctx, rc, "stq_device_alloc");
  std::vector<stq_pair> host;
  try {
    host.resize((size_t)count);
  } catch (const std::bad_alloc &) {
    stq_free_device(ctx, d_pairs);
    return false;
  }
  int64_t written = 0;
  rc = stq_find_pairs_rows_device(ctx, d.get(), N, 0, rows, static_cast<stq_pair *>(d_pairs), count,
                                  &written);
  if (rc == STQ_OK)
    rc = stq_copy_to_host(ctx, host.data(), d_pairs, sizeof(stq_pair) * (size_t)written);
  stq_free_device(ctx, d_pairs);
  if (rc == STQ_ERR_OUT_OF_MEMORY) return false;
  check(ctx, rc, "stq_find_pairs_rows_device");
  // the vector is in sweep order, so the smaller position is the earlier box
  overlaps.reserve((size_t)written);
  for (const stq_pair &p : host)
    overlaps.emplace_back(boxes[(size_t)p.first].id, boxes[(size_t)p.second].id);
  return true;
}
}  // namespace detail

// src/cpu/sweep.cpp:94-121 (with the merge of :79-92).
inline void sweep(const std::vector<Aabb> &boxes, std::vector<std::pair<int, int>> &overlaps,
                  int n) {
  if (!detail::sweep_rows(boxes, overlaps, n)) {
    std::fprintf(stderr, "sweep: the pairs do not fit in memory\n");
    std::exit(STQ_ERR_OUT_OF_MEMORY);
  }
}

// src/cpu/sweep.cpp:123-140: halves n while the pairs do not fit.
inline void run_sweep_cpu(std::vector<Aabb> &boxes, int &n,
                          std::vector<std::pair<int, int>> &finOverlaps) {
  finOverlaps.clear();
  if (n > (int)boxes.size()) n = (int)boxes.size();
  while (!detail::sweep_rows(boxes, finOverlaps, n)) {
    finOverlaps.clear();
    n /= 2;
    if (n == 0) {
      std::fprintf(stderr, "run_sweep_cpu: the pairs of one box do not fit in memory\n");
      std::exit(STQ_ERR_OUT_OF_MEMORY);
    }
  }
}

// src/cpu/sweep.cpp:142-158.
inline void sweep_cpu_single_batch(std::vector<Aabb> &boxes_batching, int &n, int N,
                                   std::vector<std::pair<int, int>> &overlaps) {
  overlaps.clear();
  if (boxes_batching.empty()) return;
  if ((int)boxes_batching.size() == N) sort_along_xaxis(boxes_batching);
  run_sweep_cpu(boxes_batching, n, overlaps);
  boxes_batching.erase(boxes_batching.begin(),
                       boxes_batching.begin() + std::min<size_t>((size_t)n, boxes_batching.size()));
  n = std::min(static_cast<int>(boxes_batching.size()), n);
}

}  // namespace stq::cpu
