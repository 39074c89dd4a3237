// narrow.hip -- the hand-off to a continuous-collision narrow phase.
//
// Orientation (src/simulation.cu:510-524): a vertex-face pair becomes
// (vertex box, face box), an edge-edge pair (smaller id, larger id).  Query
// records follow the layout of the reference's CCDData (src/memory.cuh:48-67,
// built without CCD_TOI_PER_QUERY): the four vertices of the pair at t0, then
// at t1 -- vertex-face: the vertex, then the face's three vertices;
// edge-edge: the first edge's two vertices, then the second's -- the minimum
// separation, and zeroed error bounds, tolerances and check count, which the
// narrow phase computes for itself.
#include "narrow.hpp"

namespace stqamd {

namespace {

constexpr int kT = 256;

__device__ inline int box_kind(const stq_aabb_f64 &b) {  // 1 vertex, 2 edge, 4 face
  return type_bit(b.vertex_ids[1], b.vertex_ids[2]);
}

// (a, b) -> oriented pair; false when the pair is not vertex-face or edge-edge
// or an id is out of range.
__device__ inline bool orient(const stq_aabb_f64 *__restrict__ boxes, int64_t n, int2 p, int2 &o,
                              int &ka, int &kb) {
  if (p.x < 0 || p.y < 0 || p.x >= n || p.y >= n) return false;
  ka = box_kind(boxes[p.x]);
  kb = box_kind(boxes[p.y]);
  if (ka == 1 && kb == 4) {
    o = p;
  } else if (ka == 4 && kb == 1) {
    o = make_int2(p.y, p.x);
    const int t = ka;
    ka = kb;
    kb = t;
  } else if (ka == 2 && kb == 2) {
    o = make_int2(min(p.x, p.y), max(p.x, p.y));
  } else {
    return false;
  }
  return true;
}

__global__ __launch_bounds__(kT) void k_orient(const stq_aabb_f64 *__restrict__ boxes, int64_t n,
                                               int2 *__restrict__ pairs, int64_t count,
                                               unsigned *__restrict__ err) {
  const int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (k >= count) return;
  int2 o;
  int ka, kb;
  if (orient(boxes, n, pairs[k], o, ka, kb)) pairs[k] = o;
  else atomicOr(err, 1u);
}

constexpr int kRec = sizeof(stq_ccd_query) / sizeof(double);  // 32 doubles
static_assert(sizeof(stq_ccd_query) == 256, "query record layout");
constexpr int kRow = kRec + 1;  // LDS row stride (doubles), against bank conflicts

// One wave (one block) per 64 queries: each lane assembles its record in LDS,
// then the wave writes the 64 records (16 KB, contiguous) with consecutive
// lanes on consecutive doubles.
__global__ __launch_bounds__(64) void k_ccd_queries(const stq_aabb_f64 *__restrict__ boxes, int64_t n,
                                                    const double *__restrict__ v0,
                                                    const double *__restrict__ v1, int64_t nv,
                                                    const int2 *__restrict__ pairs, int64_t count,
                                                    double ms, double *__restrict__ out,
                                                    unsigned *__restrict__ err) {
  __shared__ double rec[64][kRow];
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * 64;
  const int64_t k = base + lane;
  double *r = rec[lane];
  for (int j = 0; j < kRec; ++j) r[j] = 0.0;
  if (k < count) {
    int2 o;
    int ka, kb;
    int vert[4] = {0, 0, 0, 0};
    bool ok = orient(boxes, n, pairs[k], o, ka, kb);
    if (ok) {
      const stq_aabb_f64 &a = boxes[o.x];
      const stq_aabb_f64 &b = boxes[o.y];
      if (ka == 1) {  // vertex, face
        vert[0] = a.vertex_ids[0];
        vert[1] = b.vertex_ids[0];
        vert[2] = b.vertex_ids[1];
        vert[3] = b.vertex_ids[2];
      } else {  // edge, edge
        vert[0] = a.vertex_ids[0];
        vert[1] = a.vertex_ids[1];
        vert[2] = b.vertex_ids[0];
        vert[3] = b.vertex_ids[1];
      }
      for (int j = 0; j < 4; ++j) ok &= vert[j] >= 0 && (int64_t)vert[j] < nv;
    }
    if (ok) {
      for (int j = 0; j < 4; ++j)
        for (int c = 0; c < 3; ++c) {
          r[3 * j + c] = v0[3 * (int64_t)vert[j] + c];
          r[12 + 3 * j + c] = v1[3 * (int64_t)vert[j] + c];
        }
      r[24] = ms;
    } else {
      atomicOr(err, 1u);
    }
  }
  __syncthreads();
  const int64_t nrec = count - base < 64 ? count - base : 64;
  if (nrec <= 0) return;
  double *dst = out + base * kRec;
  for (int c = lane; c < nrec * kRec; c += 64) dst[c] = rec[c / kRec][c % kRec];
}

}  // namespace

void launch_orient_pairs(const stq_aabb_f64 *boxes, int64_t n, stq_pair *pairs, int64_t count,
                         unsigned *err, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_orient, dim3((unsigned)((count + kT - 1) / kT)), dim3(kT), 0, st, boxes, n,
                     reinterpret_cast<int2 *>(pairs), count, err);
  STQ_HIP(hipGetLastError());
}

void launch_ccd_queries(const stq_aabb_f64 *boxes, int64_t n, const double *v0, const double *v1,
                        int64_t nv, const stq_pair *pairs, int64_t count, double ms,
                        stq_ccd_query *out, unsigned *err, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_ccd_queries, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, st, boxes,
                     n, v0, v1, nv, reinterpret_cast<const int2 *>(pairs), count, ms,
                     reinterpret_cast<double *>(out), err);
  STQ_HIP(hipGetLastError());
}

}  // namespace stqamd
