// build.hip -- swept boxes of a moving triangle mesh, built on the GPU.
//
// Replaces the host construction of the reference, constructBoxes
// (src/io.cu:15-28) with addVertices / addEdges / addFaces
// (src/aabb.cu:63-174): one box per vertex, per edge and per face, in that
// order, covering the t0 and t1 positions of the simplex's vertices and
// inflated by a radius.  The reference merges per-thread vectors, so its box
// order varies from run to run; here box s is always at position s (id = s),
// and ref_id is the index within its kind, as the reference's Aabb
// constructor records it.
//
//   FP64 build (the reference compiled with CCD_USE_DOUBLE):
//     min = min(points) - r,  max = max(points) + r
//   float build (the reference's default FP32 build, values kept in FP64):
//     min = next_down((float)min(points)) - next_up((float)r)
//     max = next_up((float)max(points))   + next_up((float)r)
//   in float arithmetic, with next_up/next_down = nextafterf(x, x +/- FLT_MAX)
//   (src/aabb.cu:56-61): one ulp outward, except that +inf stays +inf going
//   down and -inf stays -inf going up.
#include <cfloat>
#include <cmath>

#include "build.hpp"

namespace stqamd {

namespace {

constexpr int kT = 256;

__device__ inline float step_toward(float x, float y) {
  if (x != x || y != y) return x + y;
  if (x == y) return y;
  if (x == 0.f) return y > 0.f ? 1.4e-45f : -1.4e-45f;  // +-denorm_min
  uint32_t u = __float_as_uint(x);
  u = ((x > 0.f) == (y > x)) ? u + 1u : u - 1u;
  return __uint_as_float(u);
}
__device__ inline float ref_next_up(float x) { return step_toward(x, __fadd_rn(x, FLT_MAX)); }
__device__ inline float ref_next_down(float x) { return step_toward(x, __fsub_rn(x, FLT_MAX)); }

__global__ __launch_bounds__(kT) void k_build_boxes(const double *__restrict__ v0,
                                                    const double *__restrict__ v1, int64_t nv,
                                                    const int32_t *__restrict__ edges, int64_t ne,
                                                    const int32_t *__restrict__ faces, int64_t nf,
                                                    double r, int float_build,
                                                    stq_aabb_f64 *__restrict__ out,
                                                    unsigned *__restrict__ err) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= nv + ne + nf) return;
  int32_t idx[3] = {0, 0, 0};
  int k;
  stq_aabb_f64 b;
  b.id = (int32_t)s;
  b.pad0 = 0;
  if (s < nv) {
    const int32_t i = (int32_t)s;
    idx[0] = i;
    k = 1;
    b.vertex_ids[0] = i;
    b.vertex_ids[1] = -i - 1;
    b.vertex_ids[2] = -i - 1;
    b.ref_id = i;
  } else if (s < nv + ne) {
    const int64_t i = s - nv;
    idx[0] = edges[2 * i];
    idx[1] = edges[2 * i + 1];
    k = 2;
    b.vertex_ids[0] = idx[0];
    b.vertex_ids[1] = idx[1];
    b.vertex_ids[2] = -idx[0] - 1;
    b.ref_id = (int32_t)i;
  } else {
    const int64_t i = s - nv - ne;
    idx[0] = faces[3 * i];
    idx[1] = faces[3 * i + 1];
    idx[2] = faces[3 * i + 2];
    k = 3;
    b.vertex_ids[0] = idx[0];
    b.vertex_ids[1] = idx[1];
    b.vertex_ids[2] = idx[2];
    b.ref_id = (int32_t)i;
  }
  bool bad = false;
  for (int j = 0; j < k; ++j) bad |= idx[j] < 0 || (int64_t)idx[j] >= nv;
  if (bad) {
    atomicOr(err, 1u);
    for (int a = 0; a < 3; ++a) b.min[a] = b.max[a] = NAN;
    out[s] = b;
    return;
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int j = 0; j < k; ++j) {
    const double *p = v0 + 3 * (int64_t)idx[j];
    const double *q = v1 + 3 * (int64_t)idx[j];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], fmin(p[a], q[a]));
      hi[a] = fmax(hi[a], fmax(p[a], q[a]));
    }
  }
  if (float_build) {
    const float rf = ref_next_up(__double2float_rn(r));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      b.min[a] = (double)__fsub_rn(ref_next_down(__double2float_rn(lo[a])), rf);
      b.max[a] = (double)__fadd_rn(ref_next_up(__double2float_rn(hi[a])), rf);
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      b.min[a] = __dsub_rn(lo[a], r);
      b.max[a] = __dadd_rn(hi[a], r);
    }
  }
  out[s] = b;
}

}  // namespace

void launch_build_boxes(const double *v0, const double *v1, int64_t nv, const int32_t *edges,
                        int64_t ne, const int32_t *faces, int64_t nf, double r, bool float_build,
                        stq_aabb_f64 *out, unsigned *err, hipStream_t st) {
  const int64_t n = nv + ne + nf;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_build_boxes, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, st, v0, v1,
                     nv, edges, ne, faces, nf, r, float_build ? 1 : 0, out, err);
  STQ_HIP(hipGetLastError());
}

}  // namespace stqamd
