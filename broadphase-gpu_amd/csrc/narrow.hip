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
//
// Layout: a first pass keeps one kind byte per box (kinds[], 1 B per box, a
// few MB that stay in L2), so orientation reads two bytes per pair instead of
// two 72-byte boxes; the query kernel reads the boxes' vertex ids only for
// the pairs it builds.  Query records leave through LDS in half-record
// (128-byte) pieces, 8.5 KB of LDS per wave.
#include <algorithm>

#include "narrow.hpp"

namespace stqamd {

namespace {

constexpr int kT = 256;

__global__ __launch_bounds__(kT) void k_box_kinds(const stq_aabb_f64 *__restrict__ boxes, int64_t n,
                                                  uint8_t *__restrict__ kinds) {
  const int64_t stride = (int64_t)gridDim.x * kT;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += stride)
    kinds[i] = (uint8_t)type_bit(boxes[i].vertex_ids[1], boxes[i].vertex_ids[2]);
}

// (a, b) -> oriented pair, with the kinds (1 vertex, 2 edge, 4 face) of its
// two boxes; false when the pair is not vertex-face or edge-edge or an id is
// out of range.
__device__ inline bool orient(const uint8_t *__restrict__ kinds, int64_t n, int2 p, int2 &o,
                              int &ka, int &kb) {
  if (p.x < 0 || p.y < 0 || p.x >= n || p.y >= n) return false;
  ka = kinds[p.x];
  kb = kinds[p.y];
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

__global__ __launch_bounds__(kT) void k_orient(const uint8_t *__restrict__ kinds, int64_t n,
                                               int2 *__restrict__ pairs, int64_t count,
                                               unsigned *__restrict__ err) {
  const int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (k >= count) return;
  int2 o;
  int ka, kb;
  if (orient(kinds, n, pairs[k], o, ka, kb)) pairs[k] = o;
  else atomicOr(err, 1u);
}

static_assert(sizeof(stq_ccd_query) == 256, "query record layout");
constexpr int kHalf = 16;        // doubles per half record (128 B, one cache line)
constexpr int kRow = kHalf + 1;  // LDS row stride (doubles), against bank conflicts

// Each thread assembles its pair's record in registers; each wave then writes
// its 64 records in two halves (first and last 128 bytes of every record):
// the half-records go through the wave's LDS rows and leave with consecutive
// lanes on consecutive doubles, so every store instruction writes four whole
// 128-byte lines.  The stores are non-temporal: the records are for another
// kernel, and caching them would evict the vertex positions being gathered.
__global__ __launch_bounds__(kT) void k_ccd_queries(const stq_aabb_f64 *__restrict__ boxes,
                                                    const uint8_t *__restrict__ kinds, int64_t n,
                                                    const double *__restrict__ v0,
                                                    const double *__restrict__ v1, int64_t nv,
                                                    const int2 *__restrict__ pairs, int64_t count,
                                                    double ms, double *__restrict__ out,
                                                    unsigned *__restrict__ err) {
  __shared__ double stage[kT / 64][64][kRow];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t wbase = (int64_t)blockIdx.x * kT + 64 * w;  // first record of this wave
  const int64_t k = wbase + lane;
  double r[24];
#pragma unroll
  for (int j = 0; j < 24; ++j) r[j] = 0.0;
  double m = 0.0;
  if (k < count) {
    int2 o;
    int ka, kb;
    bool ok = orient(kinds, n, pairs[k], o, ka, kb);
    if (ok) {
      const int *ia = boxes[o.x].vertex_ids, *ib = boxes[o.y].vertex_ids;
      int vert[4];
      if (ka == 1) {  // vertex, face: the vertex, then the face's three vertices
        vert[0] = ia[0];
        vert[1] = ib[0];
        vert[2] = ib[1];
        vert[3] = ib[2];
      } else {  // edge, edge: the first edge's two vertices, then the second's
        vert[0] = ia[0];
        vert[1] = ia[1];
        vert[2] = ib[0];
        vert[3] = ib[1];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) ok &= vert[j] >= 0 && (int64_t)vert[j] < nv;
      if (ok) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            r[3 * j + c] = v0[3 * (int64_t)vert[j] + c];
            r[12 + 3 * j + c] = v1[3 * (int64_t)vert[j] + c];
          }
        m = ms;
      }
    }
    if (!ok) atomicOr(err, 1u);
  }
  const int64_t left = count - wbase;
  const int nrec = left <= 0 ? 0 : (left < 64 ? (int)left : 64);
  double *row = stage[w][lane];
  // first half: the 16 positions of t0 and the first 4 of t1 ... 16 doubles
#pragma unroll
  for (int j = 0; j < kHalf; ++j) row[j] = r[j];
  __syncthreads();
  for (int c = lane; c < nrec * kHalf; c += 64)
    __builtin_nontemporal_store(stage[w][c / kHalf][c % kHalf], &out[(wbase + c / kHalf) * 32 + c % kHalf]);
  __syncthreads();
  // second half: the remaining 8 positions, ms, and 7 zero doubles
#pragma unroll
  for (int j = 0; j < 8; ++j) row[j] = r[kHalf + j];
  row[8] = m;
#pragma unroll
  for (int j = 9; j < kHalf; ++j) row[j] = 0.0;
  __syncthreads();
  for (int c = lane; c < nrec * kHalf; c += 64)
    __builtin_nontemporal_store(stage[w][c / kHalf][c % kHalf],
                                &out[(wbase + c / kHalf) * 32 + kHalf + c % kHalf]);
}

unsigned nblocks(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

void launch_box_kinds(const stq_aabb_f64 *boxes, int64_t n, uint8_t *kinds, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_box_kinds, dim3((unsigned)std::min<int64_t>(nblocks(n), 8192)), dim3(kT), 0,
                     st, boxes, n, kinds);
  STQ_HIP(hipGetLastError());
}

void launch_orient_pairs(const uint8_t *kinds, int64_t n, stq_pair *pairs, int64_t count,
                         unsigned *err, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_orient, dim3(nblocks(count)), dim3(kT), 0, st, kinds, n,
                     reinterpret_cast<int2 *>(pairs), count, err);
  STQ_HIP(hipGetLastError());
}

void launch_ccd_queries(const stq_aabb_f64 *boxes, const uint8_t *kinds, int64_t n,
                        const double *v0, const double *v1, int64_t nv, const stq_pair *pairs,
                        int64_t count, double ms, stq_ccd_query *out, unsigned *err,
                        hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_ccd_queries, dim3(nblocks(count)), dim3(kT), 0, st, boxes, kinds, n, v0, v1,
                     nv, reinterpret_cast<const int2 *>(pairs), count, ms,
                     reinterpret_cast<double *>(out), err);
  STQ_HIP(hipGetLastError());
}

}  // namespace stqamd
