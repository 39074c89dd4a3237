// narrow.hpp -- orientation of pairs and narrow-phase query records (narrow.hip).
#pragma once

#include "stq_internal.hpp"

namespace stqamd {

// kinds[i] = kind bit of box i (1 vertex, 2 edge, 4 face).
void launch_box_kinds(const stq_aabb_f64 *boxes, int64_t n, uint8_t *kinds, hipStream_t st);
// In place: vertex-face pairs become (vertex, face), edge-edge pairs
// (min, max).  *err |= 1 for a pair of another kind or an id outside [0, n).
void launch_orient_pairs(const uint8_t *kinds, int64_t n, stq_pair *pairs, int64_t count,
                         unsigned *err, hipStream_t st);
// One stq_ccd_query per pair (oriented as above); *err |= 1 as above or when
// a vertex id is outside [0, nv) (that record is zero).
void launch_ccd_queries(const stq_aabb_f64 *boxes, const uint8_t *kinds, int64_t n,
                        const double *v0, const double *v1, int64_t nv, const stq_pair *pairs,
                        int64_t count, double ms, stq_ccd_query *out, unsigned *err,
                        hipStream_t st);

}  // namespace stqamd
