/*
 * stq.h -- C-ABI of the MI355X (gfx950) sweep-and-tiniest-queue broad phase.
 *
 * Drop-in boundary for the reference GPU entry
 *   stq::gpu::run_sweep_sharedqueue(const Aabb *boxes, MemHandler *memhandle,
 *       int N, int nbox, std::vector<std::pair<int,int>> &finOverlaps,
 *       int2 *&d_overlaps, int *&d_count, int &threads, int &tidstart,
 *       int &devcount, const int memlimit)
 *   (reference src/simulation.cuh:22-27, body src/simulation.cu:555-782)
 * and of the reference CPU entries
 *   stq::cpu::sweep_cpu_single_batch, run_sweep_cpu, sweep, sort_along_xaxis
 *   (src/cpu/sweep.hpp:16-25), which include/stq/cpu/stq_cpu_dropin.hpp
 *   implements over this ABI (the pairs are computed on the GPU).
 *
 * Plain C types only: no HIP, CUDA or torch types cross this boundary.  Device
 * pointers are plain `void *`-compatible pointers into the current device's
 * memory.  Every function returns an int status (enum stq_status); nothing
 * calls exit() and no exception crosses the ABI (the reference exits on GPU
 * errors, src/memory.cuh:11-19).  A context owns one device ordinal, one HIP
 * stream, grow-only scratch, the autotune table and the sizing network; calls
 * on one context must be serialised by the caller, distinct contexts may be
 * used from distinct threads.  Every call is synchronous: it returns when its
 * results are complete (the reference synchronises too, src/simulation.cu:694).
 *
 * Pair semantics (identical to the reference):
 *   boxes a != b form a pair iff
 *     (1) they overlap as closed intervals on all three axes (a.min <= b.max
 *         and b.min <= a.max per axis; src/sweep.cu:245,269 and
 *         src/collision.cu:38-42),
 *     (2) is_valid_pair: (vertex,face), (face,vertex) or (edge,edge), where
 *         face <=> v[2] >= 0, edge <=> v[2] < 0 && v[1] >= 0,
 *         vertex <=> v[2] < 0 && v[1] < 0 (src/aabb.cu:12-40),
 *     (3) they share no vertex id: none of the 9 equalities a.v[s] == b.v[t]
 *         (src/collision.cu:44-48).
 *   Ids are input positions, as the reference GPU path emits
 *   (src/sweep.cu:160).  Output pairs are canonical: first < second.  The
 *   order of pairs in the output is deterministic for a given input and
 *   options; a caller that wants lexicographic order sorts them.
 *   Boxes must have no NaN and min <= max on every axis, else the call
 *   returns STQ_ERR_INVALID_BOX.
 */
#ifndef STQ_H
#define STQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The reference FP64 box, byte for byte (stq::gpu::Aabb with CCD_USE_DOUBLE,
   src/aabb.cuh:15-43): a std::vector<stq::gpu::Aabb> buffer of that build can
   be passed as-is. */
typedef struct stq_aabb_f64 {
  int32_t id;            /* offset  0 (ignored: ids are positions)        */
  int32_t pad0;          /* offset  4 (alignment padding)                 */
  double min[3];         /* offset  8                                     */
  double max[3];         /* offset 32                                     */
  int32_t vertex_ids[3]; /* offset 56: vertex {v,-v-1,-v-1}, edge {a,b,-a-1},
                            face {a,b,c}                                  */
  int32_t ref_id;        /* offset 68 (ignored)                           */
} stq_aabb_f64;          /* 72 bytes */

/* One output pair; the reference's int2 (src/collision.cu:58). */
typedef struct stq_pair {
  int32_t first;
  int32_t second;
} stq_pair;

enum stq_status {
  STQ_OK = 0,
  STQ_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative size, bad option   */
  STQ_ERR_INVALID_BOX = 2,      /* NaN, or min > max on some axis            */
  STQ_ERR_CAPACITY = 3,         /* output buffer too small; *count = needed  */
  STQ_ERR_OUT_OF_MEMORY = 4,    /* device or host allocation failed          */
  STQ_ERR_HIP = 5,              /* HIP runtime error; see stq_last_error     */
  STQ_ERR_TOO_LARGE = 6,        /* more than 2^31-1 boxes in one scene       */
  STQ_ERR_INTERNAL = 7
};

/* Arithmetic of the sweep. */
enum stq_precision {
  /* Plain FP64: sort keys and every comparison in FP64 (the reference built
     with CCD_USE_DOUBLE). */
  STQ_PRECISION_F64 = 0,
  /* Reduced-precision filters, FP64 confirmation.  The sort keys and chain
     ends use the sweep axis's FP32 bounds rounded outward (round to nearest,
     then one ulp outward: the reference rule of src/aabb.cu:56-61),
     quantised outward once more to 16-, 24- or 32-bit integers over the
     scene's range (16 bits up to 2^22 boxes per scene); the two other axes
     are tested on 14-bit levels per scene, rounded outward, with the kind
     codes of the type filter folded in.  Every filter keeps a superset of
     the FP64 candidates, and every survivor is re-tested in FP64 on all
     three axes against the original boxes, so the pair set equals the plain
     FP64 set exactly.  FP64 also computes the axis statistics and the
     quantisation ranges. */
  STQ_PRECISION_F32_CONFIRM = 1,
  /* FP64 keys and every test in FP64, with the work organisation of the
     reduced-precision form (rows grouped close on axis 1, a union test per 64
     positions, then the row tests): the ablation that separates that
     organisation from the reduced precision.  The same pair set. */
  STQ_PRECISION_F64_GROUPED = 2
};

/* Choice of the sweep axis.  The pair set never depends on it. */
enum stq_axis_policy {
  STQ_AXIS_X = 0, /* the reference's fixed choice (src/simulation.cu:637) */
  STQ_AXIS_Y = 1,
  STQ_AXIS_Z = 2,
  /* largest variance of the box end points about their mean, the rule of the
     reference's disabled calc_mean/calc_variance (src/sweep.cu:93-123,
     src/simulation.cu:594-636), computed in FP64 with a fixed-order reduction */
  STQ_AXIS_VARIANCE = 3,
  /* smallest expected chain length: argmin over axes of mean extent divided by
     the standard deviation of the box centres */
  STQ_AXIS_CHAIN = 4
};

/* How the output is sized. */
enum stq_sizing {
  /* exact count pass, prefix sum, fill pass (the reference's count-then-refill
     double pass, src/simulation.cu:451-489) */
  STQ_SIZING_COUNT_FILL = 0,
  /* one sweep pass into per-tile buffers sized by the learned per-tile pair
     predictor; tiles that overflow are redone by the exact count pass.  One
     scene: the scratch holds exactly the predicted slots.  Batched scenes:
     the buffers come from a scratch budget the context keeps (and grows)
     across calls, so a call needs no host read-back; tiles past it are redone
     too, so the result never depends on it */
  STQ_SIZING_LEARNED = 1
};

typedef struct stq_options {
  int32_t precision;     /* enum stq_precision                    */
  int32_t axis_policy;   /* enum stq_axis_policy                  */
  int32_t sizing;        /* enum stq_sizing                       */
  int32_t autotune;      /* 1: pick launch parameters by timing    */
  int32_t chunk;         /* q-steps per sweep work item; 0 = auto
                            (2048 plain FP64, 16384 reduced)       */
  int32_t window;        /* rows regrouped per tile: 0 = auto
                            (4096 reduced precision), or a power
                            of two in [64, 4096] (at most 2048 is
                            used by the plain FP64 form)           */
  int32_t waves_per_cu;  /* sweep residency per CU; 0 = auto      */
  int32_t batch;         /* q records per scalar-load batch of the
                            plain FP64 sweep: 0 = auto, or 2, 4, 8 */
  int32_t xcd_split;     /* work split over XCDs: 0 = auto (on),
                            1 = on, 2 = off                        */
  int32_t train_scenes;  /* learned sizing: scenes run count-then-fill
                            while the network trains on their exact
                            counts before the single learned pass is
                            used; 0 = 200, negative = none        */
  int32_t graph;         /* batched calls: 0 = auto (the second call
                            of an unchanged batch is captured in a
                            hipGraph, later ones replay it), 1 = capture
                            from the first call, 2 = never          */
  int32_t claim;         /* sweep work split: 0 = auto (waves claim
                            items from per-XCD counters), 1 = claim,
                            2 = fixed stride                        */
  int32_t tails;         /* long-chain rows swept as tail items with
                            positions across lanes: 0 = auto (on for
                            plain FP64, off for reduced precision),
                            1 = on, 2 = off                         */
  int32_t tail_cost;     /* weight of a tail (row, position) test
                            against a body lane step, in percent;
                            0 = 400                                 */
  int32_t scratch_limit_mb; /* learned sizing: most MiB of the pair
                            scratch (0 = no limit); items whose
                            predicted slots end beyond it are redone
                            by the exact refill, so the pairs do not
                            depend on it                           */
  int32_t balance;       /* work split over the XCD slices (with
                            xcd_split on), by cost: each item's cost
                            from its positions, rows and expected
                            pairs (the network's prediction before
                            the learned pass, the exact counts before
                            a fill pass, none before a count pass),
                            and each XCD gets an equal share of the
                            total.  0 = auto (by cost before the
                            learned pass, else equal numbers of
                            items), 1 = by cost before every pass,
                            2 = equal numbers of items             */
} stq_options;

typedef struct stq_stats {
  int64_t n_boxes;          /* boxes in the last call                       */
  int64_t n_pairs;          /* pairs in the last call                       */
  int64_t n_candidates;     /* sweep-axis candidates (sum of chain lengths) */
  int64_t n_lane_steps;     /* lane slots the sweep kernel executed         */
  int64_t n_items;          /* sweep work items                             */
  int64_t n_confirm;        /* FP32 candidates re-tested in FP64            */
  int64_t n_tiles;          /* sizing tiles                                 */
  int64_t n_fallback_tiles; /* tiles redone by the exact count pass         */
  int64_t n_scenes;         /* scenes in the last call                      */
  int32_t axis;             /* sweep axis used (last scene)                 */
  int32_t precision;
  int32_t sizing;
  int32_t graph;            /* batched call: 0 eager launches, 1 captured
                               into a graph and launched, 2 graph replay  */
  /* milliseconds, measured with HIP events on the context's stream */
  double ms_total;          /* whole call on the stream                     */
  double ms_keys;           /* validation, axis choice and key build        */
  double ms_sort;           /* radix sort                                   */
  double ms_gather;         /* sorted record build and chain ends           */
  double ms_plan;           /* work planning                                */
  double ms_sweep_count;    /* exact count sweep (0 if not run)             */
  double ms_scan;           /* prefix sums                                  */
  double ms_sweep_fill;     /* emitting sweep                               */
  double ms_compact;        /* output compaction                            */
  double ms_predict;        /* sizing network inference                     */
  double sizing_mae;        /* mean |predicted - exact| pairs per tile      */
  double ms_tune;           /* host time the autotuner spent in this call
                               (0 unless the call's shape class was new)   */
  double reserved[5];
} stq_stats;

/* Fills *opt with the defaults: plain FP64, axis x, count-then-fill. */
int stq_options_default(stq_options *opt);

/* Creates a context on HIP device `device` (replaces the reference's
   setup(), src/util.cu:9-55, and its MemHandler, src/memory.cuh:69-191).
   `opt` may be NULL for the defaults. */
int stq_create(void **ctx, int device, const stq_options *opt);
int stq_destroy(void *ctx);
int stq_set_options(void *ctx, const stq_options *opt);

/* Device-resident broad phase: boxes already in device memory, pairs written
   to the caller's device buffer `d_out` of `capacity` pairs.  On return *count
   holds the number of pairs.  If it exceeds `capacity` the call returns
   STQ_ERR_CAPACITY and the contents of d_out are unspecified.
   Replaces run_sweep_sharedqueue's d_overlaps / d_count outputs
   (src/simulation.cu:675-686, :752-753), without its capacity-driven batching
   (src/memory.cuh:171-188): the whole scene is always processed. */
int stq_find_pairs_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n,
                          stq_pair *d_out, int64_t capacity, int64_t *count);

/* ---- allocation-free form (no counterpart in the reference) ----
   For callers that cannot allocate device memory during the call (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
   the same pairs as stq_find_pairs_device, with every scratch buffer taken
   from one fixed device array per GPU (STQ_NOALLOC_SCRATCH_BYTES, reserved
   when the library is loaded on that GPU and held until the process ends).
   The call allocates no device or host memory.  Sizing is count-then-fill
   (options.sizing is ignored), autotuning only applies a table entry that
   already exists, and planning caps the work items per 64-row group, so the
   pair order may differ from stq_find_pairs_device's (the set is the same).
   More than STQ_NOALLOC_MAX_BOXES boxes: STQ_ERR_TOO_LARGE, nothing runs.
   Every call on one GPU shares that GPU's array: calls that overlap on one
   GPU (from two contexts or two threads) race and must be serialised by the
   caller. */
#define STQ_NOALLOC_MAX_BOXES 262144
#define STQ_NOALLOC_SCRATCH_BYTES (48u << 20)
int stq_find_pairs_device_noalloc(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n,
                                  stq_pair *d_out, int64_t capacity, int64_t *count);

/* ---- sorted-row ranges (the reference's tidstart batching) ----
   The pairs whose sorted-earlier box has rank in [row_begin, row_end) of the
   order the reference sweeps: FP64 minima on x, ascending, ties by input
   position (thrust::sort_by_key on the x minimum, src/simulation.cu:645-649;
   rows of twostage_queue in [tidstart, tidstart + MAX_OVERLAP_CUTOFF),
   src/sweep.cu:234-240).  The union over ranges that cover [0, n) is the
   whole pair set, each pair in exactly one range.  Runs the plain FP64 form
   with count-then-fill whatever the context's options; ids and orientation
   as stq_find_pairs_device; row_end may exceed n. */
int stq_find_pairs_rows_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n,
                               int64_t row_begin, int64_t row_end, stq_pair *d_out,
                               int64_t capacity, int64_t *count);
/* The input positions of the boxes in that order (n int32 to d_order): the
   key sort of the sweep.  STQ_ERR_INVALID_BOX as above. */
int stq_sort_order_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n, int32_t *d_order);
/* Free and total bytes of the context's device (for callers that size their
   own buffers, as the reference's MemHandler does, src/memory.cuh:78-94). */
int stq_device_memory(void *ctx, size_t *free_bytes, size_t *total_bytes);
/* The sweep's key sort (replaces thrust::sort_by_key, src/simulation.cu:
   645-649) on caller keys: a stable LSD radix sort of n (key, value) pairs
   by the low key_bits bits of 4- or 8-byte unsigned keys, in place on the
   device.  form 0: digits equal for every key are skipped (one host
   read-back of the digit histogram); 1: no host read-back (such digits cost
   a copy pass; capturable); 2: 4-byte keys, every digit a pass, no
   histogram. */
int stq_radix_sort_device(void *ctx, void *d_keys, uint32_t *d_vals, int64_t n, int32_t key_bytes,
                          int32_t key_bits, int32_t form);

/* As stq_find_pairs_device, but the library allocates *d_out with exactly
   *count pairs (at least one), like the reference's cudaMalloc'd d_overlaps
   that the caller owns (src/simulation.cu:686); release with stq_free_device. */
int stq_find_pairs_device_alloc(void *ctx, const stq_aabb_f64 *d_boxes,
                                int64_t n, stq_pair **d_out, int64_t *count);

/* Host in, host out: copies the boxes in (src/simulation.cu:578-580), runs,
   and copies the pairs back (the KEEP_CPU_OVERLAPS branch,
   src/simulation.cu:763-780).  *h_out is malloc'd; release with
   stq_free_host. */
int stq_find_pairs_host(void *ctx, const stq_aabb_f64 *h_boxes, int64_t n,
                        stq_pair **h_out, int64_t *count);

/* Batched form: `nscenes` independent scenes stored back to back in d_boxes;
   scene s holds boxes [h_scene_offsets[s], h_scene_offsets[s+1]) and its ids
   are positions within the scene.  Pairs of scene s are written to
   d_out[h_pair_offsets[s] .. h_pair_offsets[s+1]); h_pair_offsets has
   nscenes+1 entries.  STQ_ERR_CAPACITY as above (h_pair_offsets[nscenes] then
   holds the total needed).  No counterpart in the reference, which runs one
   scene per call. */
int stq_find_pairs_batched_device(void *ctx, const stq_aabb_f64 *d_boxes,
                                  const int64_t *h_scene_offsets,
                                  int32_t nscenes, stq_pair *d_out,
                                  int64_t capacity, int64_t *h_pair_offsets);

/* ---- swept boxes of a moving triangle mesh ----
   Replaces the reference's constructBoxes (src/io.cu:15-28) with addVertices,
   addEdges and addFaces (src/aabb.cu:63-174), which run on the host there.
   d_v0/d_v1: nv vertices at t0 and t1, row-major (x, y, z) doubles;
   d_edges: ne pairs of vertex indices; d_faces: nf triples.  Writes
   nv + ne + nf boxes to d_boxes: vertices, then edges, then faces; box s has
   id s and ref_id = its index within its kind; vertex ids follow the
   reference convention.  Bounds cover both positions of every vertex of the
   simplex, inflated by `inflation_radius`:
     float_build = 0: FP64 (the reference built with CCD_USE_DOUBLE);
     float_build = 1: the reference's FP32 build, bit for bit (rounded to
                      float, one ulp outward as nextafterf, radius rounded
                      up, float arithmetic), stored as doubles.
   STQ_ERR_INVALID_ARGUMENT if an edge or face index is outside [0, nv).
   After this call and the two hand-off calls below, stq_get_stats gives the
   records written in n_boxes (boxes here, pairs there) and the kernel's time
   on the context's stream in ms_total; the other fields are zero. */
int stq_build_swept_boxes_device(void *ctx, const double *d_v0, const double *d_v1, int64_t nv,
                                 const int32_t *d_edges, int64_t ne, const int32_t *d_faces,
                                 int64_t nf, double inflation_radius, int32_t float_build,
                                 stq_aabb_f64 *d_boxes);
/* The boxes above (FP64 build) straight into stq_find_pairs_device: pair ids
   are box positions (vertices, then edges, then faces).  The boxes live in
   the context's scratch. */
int stq_find_pairs_mesh_device(void *ctx, const double *d_v0, const double *d_v1, int64_t nv,
                               const int32_t *d_edges, int64_t ne, const int32_t *d_faces,
                               int64_t nf, double inflation_radius, stq_pair *d_out,
                               int64_t capacity, int64_t *count);

/* ---- the narrow-phase hand-off ----
   The reference hands its pairs to a continuous-collision narrow phase
   oriented as (vertex box, face box) for vertex-face pairs and (smaller id,
   larger id) for edge-edge pairs (src/simulation.cu:510-524), and sizes that
   narrow phase's CCDData queries (src/memory.cuh:48-67,96-161). */
typedef struct stq_ccd_query {
  double v0s[3], v1s[3], v2s[3], v3s[3]; /* the pair's four vertices at t0 */
  double v0e[3], v1e[3], v2e[3], v3e[3]; /* ... and at t1                   */
  double ms;                             /* minimum separation              */
  double err[3];                         /* zero: set by the narrow phase   */
  double tol[3];                         /* zero: set by the narrow phase   */
  int32_t nbr_checks;                    /* zero                            */
  int32_t pad;
} stq_ccd_query; /* 256 bytes: CCDData built without CCD_TOI_PER_QUERY */

/* In place, on the device: vertex-face pairs become (vertex, face), edge-edge
   pairs (min, max).  STQ_ERR_INVALID_ARGUMENT if a pair is of another kind
   or an id is outside [0, n). */
int stq_orient_pairs_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n, stq_pair *d_pairs,
                            int64_t count);
/* One query per pair (oriented as above): vertex-face: the vertex, then the
   face's three vertices; edge-edge: the first edge's two vertices, then the
   second's; positions from d_v0 (t0) and d_v1 (t1), row-major (x, y, z).
   The vertices are the boxes' vertex ids. */
int stq_ccd_queries_device(void *ctx, const stq_aabb_f64 *d_boxes, int64_t n, const double *d_v0,
                           const double *d_v1, int64_t nv, const stq_pair *d_pairs, int64_t count,
                           double min_separation, stq_ccd_query *d_queries);
int stq_sizeof_ccd_query(void);

/* ---- mesh input and ground truth (host only, no context) ----
   stq_mesh_read replaces parseMesh's igl::readOBJ / igl::readPLY
   (src/io.cu:30-51): OBJ ("v" and triangular "f" lines) or PLY (ascii or
   binary little endian, vertex x/y/z, triangular faces).  *h_vertices gets
   nv rows of (x, y, z) doubles, *h_faces nf triples of 0-based indices.
   stq_mesh_edges replaces igl::edges: every undirected edge of the faces once
   as (smaller, larger), ordered by the larger index, then the smaller.
   stq_recall replaces compare_mathematica (src/groundtruth.cu:32-67): the
   number of pairs of the JSON list [[i, j], ...] found among h_pairs, and
   the list's size (canonical != 0: both sides as (min, max)).
   Outputs are malloc'd; release them with stq_free_host. */
int stq_mesh_read(const char *path, double **h_vertices, int64_t *nv, int32_t **h_faces,
                  int64_t *nf);
int stq_mesh_edges(const int32_t *h_faces, int64_t nf, int32_t **h_edges, int64_t *ne);
int stq_recall(const stq_pair *h_pairs, int64_t count, const char *json_path, int32_t canonical,
               int64_t *true_positives, int64_t *total);

/* ---- autotuning (no counterpart in the reference, whose setup(),
   src/util.cu:9-55, derives one block size from device attributes) ----
   With options.autotune = 1, the first call of each shape class (precision,
   batched or not, floor(log2 boxes), floor(log2 scenes)) with at least 16384
   boxes runs count-only trials on its own input and picks the sweep-axis
   policy, window, chunk, batch, waves_per_cu and tail_cost (0: no tail items, so
   the tuned value also sets tails); those values then replace the options' own for
   every call of that class.  The table belongs to the
   context and can be saved to and loaded from a text file. */
int stq_autotune_save(void *ctx, const char *path);
int stq_autotune_load(void *ctx, const char *path);
int stq_autotune_clear(void *ctx);
int stq_autotune_entries(void *ctx, int32_t *count);

/* Device memory plumbing for callers that do not link the HIP runtime
   themselves (foreign-function callers, the C++ drop-in of
   include/stq/gpu/stq_dropin.hpp): allocate on the context's device, copy in
   and out on the context's stream (synchronous). */
int stq_device_alloc(void *ctx, size_t bytes, void **d_ptr);
int stq_copy_to_device(void *ctx, void *d_dst, const void *h_src, size_t bytes);
int stq_copy_to_host(void *ctx, void *h_dst, const void *d_src, size_t bytes);
int stq_free_device(void *ctx, void *d_ptr);
void stq_free_host(void *h_ptr);

/* ---- the learned sizing network (no counterpart in the reference) ----
   An MLP (32 features -> 64 -> 64 -> 1, ReLU, FP32 weights) predicts the
   log1p of each sweep work item's pair count.  Its parameters live on the
   context's device; these entry points expose it for inspection, for
   gradient checks and for offline training.  Feature rows are 32 floats;
   the parameter vector is W1[64][32], b1[64], W2[64][64], b2[64], W3[64], b3. */
int stq_sizer_num_params(void);
int stq_sizer_num_features(void);
int stq_sizer_get_params(void *ctx, float *h_params);
int stq_sizer_set_params(void *ctx, const float *h_params);
/* Re-initialises the parameters (deterministic in seed), the Adam state and
   the count of scenes seen. */
int stq_sizer_reset(void *ctx, uint64_t seed);
int stq_sizer_forward(void *ctx, const float *h_features, int64_t n, float *h_out);
/* Mean squared error over n rows and its gradient (num_params floats). */
int stq_sizer_gradient(void *ctx, const float *h_features, const float *h_targets, int64_t n,
                       float *h_grad, float *h_loss);
/* `steps` Adam steps with learning rate lr on the given rows. */
int stq_sizer_train(void *ctx, const float *h_features, const float *h_targets, int64_t n,
                    int32_t steps, float lr);
/* Features and exact pair counts of the work items of the last call made
   with learned sizing (the training data the network sees).  *n receives
   the number of items; at most `capacity` rows are copied. */
int stq_sizer_last_items(void *ctx, float *h_features, uint32_t *h_counts, int64_t capacity,
                         int64_t *n);

/* Statistics of the last call on this context. */
int stq_get_stats(void *ctx, stq_stats *stats);

/* Message of the last failing call on this context ("" if none). */
const char *stq_last_error(void *ctx);
const char *stq_status_string(int status);

/* sizeof checks for foreign-function callers. */
int stq_sizeof_aabb(void);
int stq_sizeof_options(void);
int stq_sizeof_stats(void);

#ifdef __cplusplus
}
#endif

#endif /* STQ_H */
