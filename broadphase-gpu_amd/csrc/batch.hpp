// batch.hpp -- the batched (many-scene) pipeline: segmented keys, per-scene
// statistics, segment-bounded chain ends and planning, per-scene pair offsets.
// Every size it needs on the device is either fixed by the scene offsets
// (known on the host) or read from device memory, so the whole sequence
// runs without a host synchronisation and can be captured in a hipGraph.
// No counterpart in the reference, which runs one scene per call.
#pragma once

#include <vector>

#include "stq_internal.hpp"
#include "sweep.hpp"

namespace stqamd {

// Host-built tables of one batch shape (scene offsets + window size).
struct BatchTables {
  int nscenes = 0;
  int64_t n = 0;                       // boxes in all scenes
  int64_t max_scene = 0;               // boxes in the largest scene
  std::vector<int64_t> seg_off;        // nscenes + 1
  std::vector<int> chunk_scene;        // statistics chunks: scene, [start, end)
  std::vector<int64_t> chunk_start, chunk_end;
  std::vector<int> scene_chunk_begin;  // nscenes + 1
  std::vector<int64_t> win_start;      // planning windows (never straddle scenes)
  std::vector<int> win_cnt;
  std::vector<int64_t> win_gbase;      // first group of each window
  std::vector<int64_t> scene_first_group;  // nscenes + 1 (a group id, or ngroups)
  int64_t ngroups = 0;
  std::vector<SegTile> sort_tiles;     // tiles of the segmented key sort
};
void build_batch_tables(const int64_t *h_scene_offsets, int nscenes, int W, BatchTables &t);

// Device copies of the tables.
struct BatchDev {
  int64_t *seg_off = nullptr;
  int *chunk_scene = nullptr;
  int64_t *chunk_start = nullptr, *chunk_end = nullptr;
  int *scene_chunk_begin = nullptr;
  int64_t *win_start = nullptr;
  int *win_cnt = nullptr;
  int64_t *win_gbase = nullptr;
  int64_t *scene_first_group = nullptr;
  SceneParams *params = nullptr;
  double *partial = nullptr;  // nchunks * kAxisStats
  SegTile *sort_tiles = nullptr;
};

// keys3 (reduced precision; null: none): the FP32 sort keys of all three
// axes (3 n words), written by the statistics pass itself, which then also
// sets kErrInverted in *err for NaN or inverted boxes; the scene parameters
// carry the quantisation of the sweep keys, sweep_key_bits(t.max_scene) bits.
void launch_seg_stats(const stq_aabb_f64 *boxes, const BatchTables &t, const BatchDev &d,
                      int policy, int fixed_axis, uint32_t *keys3, unsigned *err, hipStream_t st);
// F32 form: 4-byte keys, the image of the rounded-down minimum, sorted per
// scene by the segmented sort; F64 form: 8-byte keys, the image of the FP64
// minimum (a second, stable sort by scene follows).
void launch_keys_seg(bool f32, const stq_aabb_f64 *boxes, const BatchTables &t, const BatchDev &d,
                     void *keys, uint32_t *vals, unsigned *err, hipStream_t st);
// scene_key[s] = scene of input position vals[s] (for the F64 second sort)
void launch_scene_keys(const uint32_t *vals, const BatchTables &t, const BatchDev &d,
                       uint32_t *scene_key, hipStream_t st);
// Sorted records of every scene; with cols and rowf, the column statistics
// of every kColBlock global positions and the row features of every position
// as well (null: not made).
void launch_gather_seg(bool f32, const stq_aabb_f64 *boxes, const uint32_t *order,
                       const BatchTables &t, const BatchDev &d, void *smin, void *smax, void *rec,
                       QRec *qrec, uint32_t *orig_local, ColStats *cols, uint2 *rowf,
                       hipStream_t st);
void launch_chain_end_seg(bool f32, const void *smin, const void *smax, const BatchTables &t,
                          const BatchDev &d, int *hi, unsigned long long *cand, hipStream_t st);
// Groups of 64 rows per window (ordered as in launch_plan: by chain end, or
// with qrec by group_key, with tails each group's rows then by chain end), planned as
// body + tail (plan.hpp); each part is cut into at most 64 items, so the item
// buffer has a bound the host computes from the scene sizes.  Items are then
// written by launch_items (sweep.hpp).
void launch_plan_seg(const int *hi, const QRec *qrec, const BatchTables &t, const BatchDev &d,
                     int W, int chunk, float tail_cost, int2 *rowlist, int4 *groups,
                     int4 *tgroups, int2 *gchunk, uint32_t *nchunks,
                     unsigned long long *lane_steps, hipStream_t st,
                     const uint2 *rowf = nullptr, uint2 *rowfeat = nullptr);
// pair_off[s] = off[ioff[scene_first_group[s]]] for s = 0..nscenes.
void launch_scene_pairs(const int64_t *off, const int64_t *ioff, const BatchTables &t,
                        const BatchDev &d, int64_t *pair_off, hipStream_t st);
void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t st);
void launch_fill_u64(unsigned long long *p, unsigned long long v, int64_t n, hipStream_t st);

}  // namespace stqamd
