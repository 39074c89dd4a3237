// build.hpp -- swept boxes of a moving triangle mesh (build.hip).
#pragma once

#include "stq_internal.hpp"

namespace stqamd {

// Boxes of nv vertices, then ne edges, then nf faces into out[0 .. nv+ne+nf).
// v0/v1: row-major (x, y, z) per vertex at t0 and t1; edges: 2 ints each;
// faces: 3 ints each.  *err |= 1 when an edge or face index is outside
// [0, nv) (that box gets NaN bounds).
void launch_build_boxes(const double *v0, const double *v1, int64_t nv, const int32_t *edges,
                        int64_t ne, const int32_t *faces, int64_t nf, double r, bool float_build,
                        stq_aabb_f64 *out, unsigned *err, hipStream_t st);

}  // namespace stqamd
