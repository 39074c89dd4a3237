/*
 * stq_oracle.c -- CPU oracle for the sweep-and-tiniest-queue broad phase.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product library (libstq_amd.so) never links or calls it.
 *
 * It restates, in plain C with OpenMP, the reference CPU broad phase of
 * Ahdhn/broadphase-gpu in FP64 (CCD_USE_DOUBLE semantics):
 *
 *   predicates      src/cpu/sweep.cpp:17-44   (is_face/is_edge/is_vertex,
 *                   is_valid_pair, does_collide on the two non-sweep axes,
 *                   covertex with its 9 equalities); identical to the GPU
 *                   predicates in src/aabb.cu:12-40 and src/collision.cu:38-48
 *   sort            src/cpu/sweep.cpp:72-74   (parallel sort by the sweep-axis
 *                   minimum)
 *   chain scan      src/cpu/sweep.cpp:94-121  (for each i, scan j = i+1.. and
 *                   break on a.max < b.min; emit (a.id, b.id))
 *   merge           src/cpu/sweep.cpp:79-92   (concatenate per-thread lists)
 *   FP32 rounding   src/aabb.cu:56-61, src/cpu/aabb.cpp:16-21 (round to
 *                   nearest float, then one ulp outward with nextafterf)
 *
 * Ids: the reference GPU path emits input positions (src/sweep.cu:160); the
 * CPU path emits Aabb::id, which equals the position for every box the
 * reference constructs (src/cpu/aabb.cpp:33,62,84).  This oracle emits input
 * positions.  Ties among equal sweep-axis minima are broken by input position
 * (a stable sort); the pair set does not depend on that choice for boxes with
 * min <= max on every axis.
 *
 * Pinning (tests/test_oracle.py): the sweep is checked against the O(N^2)
 * brute force below on random scenes of every kind, against hand-built
 * known-answer fixtures in tests/golden/, and for invariance of the set under
 * the choice of sweep axis; the rounding rule is checked against libm's
 * nextafterf and against its closed-form bounds.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct oracle_aabb {
  int32_t id;
  int32_t pad0;
  double min[3];
  double max[3];
  int32_t vids[3];
  int32_t ref_id;
} oracle_aabb; /* 72 bytes: the reference FP64 Aabb layout (src/aabb.cuh:15-43) */

/* ---- predicates: src/cpu/sweep.cpp:17-44 ---------------------------------- */
static inline int is_face(const int32_t *v) { return v[2] >= 0; }
static inline int is_edge(const int32_t *v) { return v[2] < 0 && v[1] >= 0; }
static inline int is_vertex(const int32_t *v) { return v[2] < 0 && v[1] < 0; }
static inline int is_valid_pair(const int32_t *a, const int32_t *b) {
  return (is_vertex(a) && is_face(b)) || (is_face(a) && is_vertex(b)) ||
         (is_edge(a) && is_edge(b));
}
static inline int covertex(const int32_t *a, const int32_t *b) {
  return a[0] == b[0] || a[0] == b[1] || a[0] == b[2] || a[1] == b[0] ||
         a[1] == b[1] || a[1] == b[2] || a[2] == b[0] || a[2] == b[1] ||
         a[2] == b[2];
}
/* closed overlap on the two axes other than `axis` (src/cpu/sweep.cpp:32-37) */
static inline int collide_other(const oracle_aabb *a, const oracle_aabb *b,
                                int axis) {
  int k1 = axis == 0 ? 1 : 0, k2 = axis == 2 ? 1 : 2;
  return a->max[k1] >= b->min[k1] && a->min[k1] <= b->max[k1] &&
         a->max[k2] >= b->min[k2] && a->min[k2] <= b->max[k2];
}

int oracle_sizeof_aabb(void) { return (int)sizeof(oracle_aabb); }

int oracle_pair_predicate(const oracle_aabb *a, const oracle_aabb *b) {
  /* full FP64 test of one unordered pair: closed overlap on all three axes,
     type filter, no shared vertex. */
  for (int k = 0; k < 3; ++k)
    if (!(a->max[k] >= b->min[k] && a->min[k] <= b->max[k])) return 0;
  return is_valid_pair(a->vids, b->vids) && !covertex(a->vids, b->vids);
}

/* ---- FP32 outward rounding: src/aabb.cu:56-61 ------------------------------ */
float oracle_round_down(double d) {
  float f = (float)d;
  return nextafterf(f, f - 3.40282347e+38f);
}
float oracle_round_up(double d) {
  float f = (float)d;
  return nextafterf(f, f + 3.40282347e+38f);
}
void oracle_round_outward(const double *in, int64_t n, float *lo, float *hi) {
  for (int64_t i = 0; i < n; ++i) {
    lo[i] = oracle_round_down(in[i]);
    hi[i] = oracle_round_up(in[i]);
  }
}
/* Boxes whose coordinates are the FP32 outward-rounded values (stored back in
   doubles, exactly).  Sweeping these reproduces the reference default build
   (Scalar = float), whose set is a superset of the FP64 set. */
void oracle_boxes_round_outward(const oracle_aabb *in, int64_t n,
                                oracle_aabb *out) {
  for (int64_t i = 0; i < n; ++i) {
    out[i] = in[i];
    for (int k = 0; k < 3; ++k) {
      out[i].min[k] = (double)oracle_round_down(in[i].min[k]);
      out[i].max[k] = (double)oracle_round_up(in[i].max[k]);
    }
  }
}

/* ---- sort by sweep-axis minimum (src/cpu/sweep.cpp:72-74) ------------------
   Parallel merge sort on (key, position); ties broken by position. */
typedef struct { double key; int32_t pos; } kp_t;

static inline int kp_less(const kp_t *a, const kp_t *b) {
  return a->key < b->key || (a->key == b->key && a->pos < b->pos);
}
static int kp_cmp(const void *x, const void *y) {
  const kp_t *a = (const kp_t *)x, *b = (const kp_t *)y;
  if (kp_less(a, b)) return -1;
  if (kp_less(b, a)) return 1;
  return 0;
}
static void merge_into(const kp_t *a, int64_t na, const kp_t *b, int64_t nb,
                       kp_t *out) {
  int64_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) out[k++] = kp_less(&b[j], &a[i]) ? b[j++] : a[i++];
  while (i < na) out[k++] = a[i++];
  while (j < nb) out[k++] = b[j++];
}
static void msort(kp_t *a, kp_t *tmp, int64_t n, int depth) {
  if (n < 65536 || depth <= 0) {
    qsort(a, (size_t)n, sizeof(kp_t), kp_cmp);
    return;
  }
  int64_t h = n / 2;
#pragma omp task shared(a, tmp)
  msort(a, tmp, h, depth - 1);
#pragma omp task shared(a, tmp)
  msort(a + h, tmp + h, n - h, depth - 1);
#pragma omp taskwait
  merge_into(a, h, a + h, n - h, tmp);
  memcpy(a, tmp, (size_t)n * sizeof(kp_t));
}
static void parallel_sort_kp(kp_t *a, int64_t n, int nthreads) {
  kp_t *tmp = (kp_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(kp_t));
  int depth = 0;
  while ((1 << depth) < 4 * nthreads && depth < 12) ++depth;
#pragma omp parallel num_threads(nthreads)
#pragma omp single
  msort(a, tmp, n, depth);
  free(tmp);
}

/* ---- the sweep (src/cpu/sweep.cpp:94-121, merge :79-92) --------------------
   Sorts by min[axis], then scans rows i in [0, nrows) of the sorted order
   (nrows = n for the whole scene; the reference's `n` batching argument).
   Returns the number of pairs; *out receives malloc'd (earlier, later) id
   pairs in sorted-row order.  Ids are input positions. */
typedef struct { int32_t *v; int64_t n, cap; } vec_t;
static void vpush(vec_t *v, int32_t a, int32_t b) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 1024;
    v->v = (int32_t *)realloc(v->v, (size_t)v->cap * 2 * sizeof(int32_t));
  }
  v->v[2 * v->n] = a;
  v->v[2 * v->n + 1] = b;
  v->n++;
}

int64_t oracle_sweep_rows(const oracle_aabb *boxes, int64_t n, int axis,
                          int64_t nrows, int nthreads, int32_t **out,
                          double *sort_seconds) {
  if (nthreads < 1) nthreads = 1;
  if (nrows > n || nrows < 0) nrows = n;
  double t0 = omp_get_wtime();
  kp_t *kp = (kp_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(kp_t));
  for (int64_t i = 0; i < n; ++i) {
    kp[i].key = boxes[i].min[axis];
    kp[i].pos = (int32_t)i;
  }
  parallel_sort_kp(kp, n, nthreads);
  /* sorted copy of the boxes, as the reference sorts the box vector itself */
  oracle_aabb *s = (oracle_aabb *)malloc((size_t)(n > 0 ? n : 1) * sizeof(oracle_aabb));
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    s[i] = boxes[kp[i].pos];
    s[i].id = kp[i].pos;
  }
  free(kp);
  if (sort_seconds) *sort_seconds = omp_get_wtime() - t0;

  vec_t *local = (vec_t *)calloc((size_t)nthreads, sizeof(vec_t));
#pragma omp parallel num_threads(nthreads)
  {
    vec_t *mine = &local[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < nrows; ++i) {
      const oracle_aabb *a = &s[i];
      for (int64_t j = i + 1; j < n; ++j) {
        const oracle_aabb *b = &s[j];
        if (a->max[axis] < b->min[axis]) break;
        if (collide_other(a, b, axis) && is_valid_pair(a->vids, b->vids) &&
            !covertex(a->vids, b->vids))
          vpush(mine, a->id, b->id);
      }
    }
  }
  int64_t total = 0;
  for (int t = 0; t < nthreads; ++t) total += local[t].n;
  int32_t *res = (int32_t *)malloc((size_t)(total > 0 ? total : 1) * 2 * sizeof(int32_t));
  int64_t off = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (local[t].n)
      memcpy(res + 2 * off, local[t].v, (size_t)local[t].n * 2 * sizeof(int32_t));
    off += local[t].n;
    free(local[t].v);
  }
  free(local);
  free(s);
  *out = res;
  return total;
}

int64_t oracle_sweep(const oracle_aabb *boxes, int64_t n, int axis,
                     int nthreads, int32_t **out) {
  return oracle_sweep_rows(boxes, n, axis, n, nthreads, out, NULL);
}

/* The whole-scene sweep with its wall time (sort, sweep and merge: the scope
   the reference CLI times, app/main.cpp:138-153) in *seconds. */
int64_t oracle_sweep_timed(const oracle_aabb *boxes, int64_t n, int axis,
                           int nthreads, int32_t **out, double *seconds) {
  double t0 = omp_get_wtime();
  int64_t c = oracle_sweep_rows(boxes, n, axis, n, nthreads, out, NULL);
  *seconds = omp_get_wtime() - t0;
  return c;
}

/* Number of sweep candidates per sorted row on an axis: the length of the
   chain the reference walks before its break (src/cpu/sweep.cpp:102-108).
   Returns the total; rows with a longer chain than *max_chain update it. */
int64_t oracle_candidates(const oracle_aabb *boxes, int64_t n, int axis,
                          int nthreads, int64_t *max_chain) {
  kp_t *kp = (kp_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(kp_t));
  double *mx = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int64_t i = 0; i < n; ++i) {
    kp[i].key = boxes[i].min[axis];
    kp[i].pos = (int32_t)i;
  }
  parallel_sort_kp(kp, n, nthreads);
  for (int64_t i = 0; i < n; ++i) mx[i] = boxes[kp[i].pos].max[axis];
  int64_t total = 0, mc = 0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : total) reduction(max : mc)
  for (int64_t i = 0; i < n; ++i) {
    /* upper bound of mx[i] among sorted keys, searched forward from i */
    int64_t lo = i, hi = n; /* invariant: key[lo] <= mx (or lo == i), key[hi] > mx */
    while (hi - lo > 1) {
      int64_t mid = lo + (hi - lo) / 2;
      if (kp[mid].key <= mx[i]) lo = mid; else hi = mid;
    }
    int64_t len = lo - i;
    total += len;
    if (len > mc) mc = len;
  }
  free(kp);
  free(mx);
  if (max_chain) *max_chain = mc;
  return total;
}

/* ---- brute force: every unordered pair, O(N^2) ----------------------------- */
int64_t oracle_bruteforce(const oracle_aabb *boxes, int64_t n, int nthreads,
                          int32_t **out) {
  if (nthreads < 1) nthreads = 1;
  vec_t *local = (vec_t *)calloc((size_t)nthreads, sizeof(vec_t));
#pragma omp parallel num_threads(nthreads)
  {
    vec_t *mine = &local[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = i + 1; j < n; ++j)
        if (oracle_pair_predicate(&boxes[i], &boxes[j]))
          vpush(mine, (int32_t)i, (int32_t)j);
  }
  int64_t total = 0;
  for (int t = 0; t < nthreads; ++t) total += local[t].n;
  int32_t *res = (int32_t *)malloc((size_t)(total > 0 ? total : 1) * 2 * sizeof(int32_t));
  int64_t off = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (local[t].n)
      memcpy(res + 2 * off, local[t].v, (size_t)local[t].n * 2 * sizeof(int32_t));
    off += local[t].n;
    free(local[t].v);
  }
  free(local);
  *out = res;
  return total;
}

/* ---- canonical form: (min,max) orientation, lexicographic sort, unique -----
   Each pair becomes the 64-bit key (min, max), each half with its sign bit
   flipped so unsigned order is the signed lexicographic order; the keys are
   sorted by an LSD radix sort (four 16-bit digits, per-thread histograms) and
   written back without repeats.  Pinned against numpy's sort in
   tests/test_oracle.py. */
static inline uint64_t pair_key(int32_t a, int32_t b) {
  const int32_t lo = a < b ? a : b, hi = a < b ? b : a;
  return ((uint64_t)((uint32_t)lo ^ 0x80000000u) << 32) | (uint64_t)((uint32_t)hi ^ 0x80000000u);
}
void oracle_sort_u64(uint64_t *keys, int64_t n, int nthreads) {
  if (n < 2) return;
  if (nthreads < 1) nthreads = 1;
  uint64_t *tmp = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
  int64_t *hist = (int64_t *)malloc((size_t)nthreads * 65536 * sizeof(int64_t));
  uint64_t *src = keys, *dst = tmp;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 16 * pass;
#pragma omp parallel num_threads(nthreads)
    {
      const int t = omp_get_thread_num(), nt = omp_get_num_threads();
      int64_t *h = hist + (size_t)t * 65536;
      memset(h, 0, 65536 * sizeof(int64_t));
      const int64_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
      for (int64_t i = i0; i < i1; ++i) h[(src[i] >> shift) & 0xffff]++;
#pragma omp barrier
#pragma omp single
      {
        /* digit-major, thread-minor exclusive offsets: a stable scatter */
        int64_t acc = 0;
        for (int d = 0; d < 65536; ++d)
          for (int u = 0; u < nt; ++u) {
            const int64_t c = hist[(size_t)u * 65536 + d];
            hist[(size_t)u * 65536 + d] = acc;
            acc += c;
          }
      }
      for (int64_t i = i0; i < i1; ++i) dst[h[(src[i] >> shift) & 0xffff]++] = src[i];
    }
    uint64_t *x = src;
    src = dst;
    dst = x;
  }
  /* four passes: the sorted keys are back in `keys` */
  free(tmp);
  free(hist);
}
int64_t oracle_canonicalize(int32_t *pairs, int64_t n) {
  if (n <= 0) return 0;
  int nthreads = omp_get_max_threads();
  if (nthreads > 64) nthreads = 64;
  uint64_t *k = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) k[i] = pair_key(pairs[2 * i], pairs[2 * i + 1]);
  oracle_sort_u64(k, n, n < 65536 ? 1 : nthreads);
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (m > 0 && k[m - 1] == k[i]) continue;
    k[m++] = k[i];
  }
  for (int64_t i = 0; i < m; ++i) {
    pairs[2 * i] = (int32_t)((uint32_t)(k[i] >> 32) ^ 0x80000000u);
    pairs[2 * i + 1] = (int32_t)((uint32_t)k[i] ^ 0x80000000u);
  }
  free(k);
  return m;
}

void oracle_free(void *p) { free(p); }

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* ---- swept-box construction: src/aabb.cu:56-174, src/io.cu:15-28 ------------
 * constructBoxes appends vertex boxes, then edge boxes, then face boxes
 * (io.cu:25-27).  Box i of a kind covers the t0 and t1 positions of its
 * simplex's vertices (2, 4 or 6 points), inflated by r:
 *   FP64 build (CCD_USE_DOUBLE):  min = min(points) - r, max = max(points) + r
 *   float build:  min = nextafter_down((float)min(points)) - nextafter_up((float)r)
 *                 max = nextafter_up((float)max(points))  + nextafter_up((float)r)
 *                 in float arithmetic, nextafter_up/down as in src/aabb.cu:56-61
 * vertexIds: vertex {i, -i-1, -i-1} (aabb.cu:113), edge {e0, e1, -e0-1}
 * (aabb.cu:80), face {f0, f1, f2} (aabb.cu:157); id = position of the box in
 * the output, ref_id = i (the index within its kind; aabb.cu:93,127,169).
 * Vertices are row-major (x, y, z) per vertex.  Returns 0, or -1 when an edge
 * or face index is outside [0, nv). */
static float ref_next_up(float x) { return nextafterf(x, x + 3.40282347e+38f); }
static float ref_next_down(float x) { return nextafterf(x, x - 3.40282347e+38f); }

static void hull_box(const double *v0, const double *v1, const int32_t *idx, int k, double r,
                     int float_build, double *mn, double *mx) {
  for (int a = 0; a < 3; ++a) {
    double lo = INFINITY, hi = -INFINITY;
    for (int j = 0; j < k; ++j) {
      const double p = v0[3 * (int64_t)idx[j] + a], q = v1[3 * (int64_t)idx[j] + a];
      lo = fmin(lo, fmin(p, q));
      hi = fmax(hi, fmax(p, q));
    }
    if (float_build) {
      const float rf = ref_next_up((float)r);
      volatile float l = ref_next_down((float)lo) - rf; /* float arithmetic */
      volatile float h = ref_next_up((float)hi) + rf;
      mn[a] = (double)l;
      mx[a] = (double)h;
    } else {
      mn[a] = lo - r;
      mx[a] = hi + r;
    }
  }
}

int oracle_build_swept_boxes(const double *v0, const double *v1, int64_t nv, const int32_t *edges,
                             int64_t ne, const int32_t *faces, int64_t nf, double r,
                             int float_build, oracle_aabb *out) {
  for (int64_t i = 0; i < ne; ++i)
    for (int j = 0; j < 2; ++j)
      if (edges[2 * i + j] < 0 || edges[2 * i + j] >= nv) return -1;
  for (int64_t i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j)
      if (faces[3 * i + j] < 0 || faces[3 * i + j] >= nv) return -1;
  const int64_t n = nv + ne + nf;
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < n; ++s) {
    oracle_aabb *b = &out[s];
    memset(b, 0, sizeof(*b));
    b->id = (int32_t)s;
    int32_t idx[3];
    int k;
    if (s < nv) {
      const int32_t i = (int32_t)s;
      idx[0] = i;
      k = 1;
      b->vids[0] = i;
      b->vids[1] = -i - 1;
      b->vids[2] = -i - 1;
      b->ref_id = i;
    } else if (s < nv + ne) {
      const int64_t i = s - nv;
      idx[0] = edges[2 * i];
      idx[1] = edges[2 * i + 1];
      k = 2;
      b->vids[0] = idx[0];
      b->vids[1] = idx[1];
      b->vids[2] = -idx[0] - 1;
      b->ref_id = (int32_t)i;
    } else {
      const int64_t i = s - nv - ne;
      for (int j = 0; j < 3; ++j) idx[j] = faces[3 * i + j];
      k = 3;
      for (int j = 0; j < 3; ++j) b->vids[j] = idx[j];
      b->ref_id = (int32_t)i;
    }
    hull_box(v0, v1, idx, k, r, float_build, b->min, b->max);
  }
  return 0;
}

/* ---- narrow-phase orientation: src/simulation.cu:510-524 --------------------
 * vertex a, face b -> (a, b); face a, vertex b -> (b, a); edge, edge ->
 * (min, max); any other pair is dropped by the reference (here: -1). */
int oracle_orient_pairs(const oracle_aabb *boxes, int64_t n, int32_t *pairs, int64_t count) {
  for (int64_t k = 0; k < count; ++k) {
    const int32_t a = pairs[2 * k], b = pairs[2 * k + 1];
    if (a < 0 || b < 0 || a >= n || b >= n) return -1;
    const int32_t *va = boxes[a].vids, *vb = boxes[b].vids;
    if (is_vertex(va) && is_face(vb)) {
      /* unchanged */
    } else if (is_edge(va) && is_edge(vb)) {
      pairs[2 * k] = a < b ? a : b;
      pairs[2 * k + 1] = a < b ? b : a;
    } else if (is_face(va) && is_vertex(vb)) {
      pairs[2 * k] = b;
      pairs[2 * k + 1] = a;
    } else {
      return -1;
    }
  }
  return 0;
}
