// tune.cpp -- the autotuner.  For each call shape class (precision, batched
// or not, floor(log2 boxes), floor(log2 scenes)) it picks the sweep-axis
// policy, the planning window, the item chunk, the scalar batch of the FP64
// sweep loop (how many q records are in flight per wave), the resident waves
// per CU and the weight of tail tests in the group cut (0: no tails), by
// timing count-only trial runs on the first input of that class:
// coordinate descent, each candidate timed twice with HIP events, a change
// kept only when it is at least 2% faster.  For a caller with learned sizing
// each trial is charged the per-item cost of the learned stages as well
// (kLearnedItemMs per item), which count-only runs do not pay.  The table can
// be saved to and loaded from a text file.
//
// Replaces the reference's setup() (src/util.cu:9-55), which derives one
// block size and shared-memory size from device attributes.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "context.hpp"

namespace {
// Per-item time of the learned pass's item-proportional stages, measured on
// the 3M-box cloth: 93k -> 286k items (chunks of 16384 -> 4096 positions)
// added 0.1 ms of features, network, caps and compaction.
constexpr double kLearnedItemMs = 5e-7;
}  // namespace
#include "plan.hpp"

namespace stqamd {

namespace {
constexpr int64_t kTuneMinBoxes = 1 << 14;  // smaller calls are not worth tuning
constexpr double kKeep = 0.98;              // a candidate must win by 2%

int floor_log2(int64_t v) { return v > 0 ? 63 - __builtin_clzll((unsigned long long)v) : 0; }
}  // namespace

uint64_t tune_key(int precision, bool batched, int64_t n, int nscenes) {
  return (uint64_t)(precision & 15) | ((uint64_t)(batched ? 1 : 0) << 4) |
         ((uint64_t)floor_log2(n) << 8) | ((uint64_t)floor_log2(nscenes) << 16);
}

void Context::autotune_and_apply(const stq_aabb_f64 *boxes, int64_t n, const int64_t *off,
                                 int nscenes, bool batched, bool trials) {
  last_tune_ms = 0;
  if (!opt.autotune) return;
  const uint64_t key = tune_key(opt.precision, batched, n, nscenes);
  auto it = tune_table.find(key);
  if (it == tune_table.end()) {
    if (n < kTuneMinBoxes || !trials) return;
    const auto t0 = std::chrono::steady_clock::now();
    const stq_options user = opt;
    // the forms with the axis-1 grouped planner (reduced precision, grouped FP64)
    const bool grouped = user.precision != STQ_PRECISION_F64;
    std::vector<int64_t> pair_off(batched ? (size_t)nscenes + 1 : 0);
    auto trial = [&](const TuneEntry &e) {
      opt = user;
      opt.sizing = STQ_SIZING_COUNT_FILL;
      opt.graph = 2;
      opt.axis_policy = e.axis_policy;
      opt.window = e.window;
      opt.chunk = e.chunk;
      opt.batch = e.batch;
      opt.waves_per_cu = e.waves_per_cu;
      opt.tails = e.tail_cost > 0 ? 1 : 2;
      opt.tail_cost = e.tail_cost;
      double best = INFINITY;
      for (int rep = 0; rep < 2; ++rep) {
        try {  // capacity 0: the count pass runs, no pair is written
          int64_t cnt = 0;
          if (batched)
            run_batched(boxes, off, nscenes, nullptr, 0, pair_off.data());
          else
            run(boxes, n, nullptr, 0, nullptr, &cnt);
        } catch (const StatusError &err) {
          if (err.status != STQ_ERR_CAPACITY) {
            opt = user;
            throw;
          }
        }
        // learned sizing adds per-item work the count-only trial does not do
        // (features, network, caps, compaction bookkeeping)
        const double extra = user.sizing == STQ_SIZING_LEARNED ? kLearnedItemMs * stats.n_items : 0.0;
        best = std::min(best, stats.ms_total + extra);
      }
      return best;
    };
    TuneEntry best{STQ_AXIS_CHAIN, batched ? (grouped ? kBatchQWindow : kBatchWindow) : 0,
                   grouped ? kQChunk : 2048, 2, 32, grouped ? 0 : 400, 0};
    best.ms = trial(best);
    auto consider = [&](TuneEntry e) {
      if (std::memcmp(&e, &best, offsetof(TuneEntry, ms)) == 0) return;
      e.ms = trial(e);
      if (e.ms < kKeep * best.ms) best = e;
    };
    for (int a : {STQ_AXIS_VARIANCE, STQ_AXIS_X}) {
      TuneEntry e = best;
      e.axis_policy = a;
      consider(e);
    }
    for (int w : {0, 256, 512, 1024, 2048, 4096}) {
      if (batched && w == 0) continue;  // no automatic window in the batched form
      if (!grouped && w > kMaxWindow64) continue;
      TuneEntry e = best;
      e.window = w;
      consider(e);
    }
    for (int c : grouped ? std::initializer_list<int>{4096, 65536} : std::initializer_list<int>{1024, 4096}) {
      TuneEntry e = best;
      e.chunk = c;
      consider(e);
    }
    // the scalar batch applies to the plain FP64 sweep only
    if (!grouped) {
      TuneEntry e = best;
      e.batch = 4;
      consider(e);
    }
    for (int w : {16, 64}) {
      TuneEntry e = best;
      e.waves_per_cu = w;
      consider(e);
    }
    for (int tc : grouped ? std::initializer_list<int>{400} : std::initializer_list<int>{200, 800}) {
      TuneEntry e = best;
      e.tail_cost = tc;
      consider(e);
    }
    opt = user;
    it = tune_table.emplace(key, best).first;
    last_tune_ms = std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - t0).count();
  }
  const TuneEntry &e = it->second;
  opt.axis_policy = e.axis_policy;
  opt.window = e.window;
  opt.chunk = e.chunk;
  opt.batch = e.batch;
  opt.waves_per_cu = e.waves_per_cu;
  opt.tails = e.tail_cost > 0 ? 1 : 2;
  opt.tail_cost = e.tail_cost;
}

// File format: a header line, then one line per entry:
//   <key hex> <axis_policy> <window> <chunk> <batch> <waves_per_cu> <tail_cost> <ms>
// (tail_cost 0: no tail items)
int Context::tune_save(const char *path) const {
  FILE *f = std::fopen(path, "w");
  if (!f) return STQ_ERR_INVALID_ARGUMENT;
  std::fprintf(f, "stq-autotune 2\n");
  for (const auto &kv : tune_table) {
    const TuneEntry &e = kv.second;
    std::fprintf(f, "%llx %d %d %d %d %d %d %.6f\n", (unsigned long long)kv.first, e.axis_policy,
                 e.window, e.chunk, e.batch, e.waves_per_cu, e.tail_cost, e.ms);
  }
  return std::fclose(f) == 0 ? STQ_OK : STQ_ERR_INVALID_ARGUMENT;
}

int Context::tune_load(const char *path) {
  FILE *f = std::fopen(path, "r");
  if (!f) return STQ_ERR_INVALID_ARGUMENT;
  char head[64] = {0};
  int version = 0;
  if (std::fscanf(f, "%63s %d", head, &version) != 2 || std::strcmp(head, "stq-autotune") != 0 ||
      version != 2) {
    std::fclose(f);
    return STQ_ERR_INVALID_ARGUMENT;
  }
  std::map<uint64_t, TuneEntry> table;
  unsigned long long key;
  TuneEntry e;
  int r;
  while ((r = std::fscanf(f, "%llx %d %d %d %d %d %d %lf", &key, &e.axis_policy, &e.window,
                          &e.chunk, &e.batch, &e.waves_per_cu, &e.tail_cost, &e.ms)) == 8) {
    const bool ok = e.axis_policy >= 0 && e.axis_policy <= 4 && e.chunk > 0 &&
                    (e.window == 0 || (e.window >= 64 && e.window <= kMaxWindow32 &&
                                       (e.window & (e.window - 1)) == 0)) &&
                    (e.batch == 2 || e.batch == 4 || e.batch == 8) && e.waves_per_cu > 0 &&
                    e.waves_per_cu <= 64 && e.tail_cost >= 0 && e.tail_cost <= 100000;
    if (!ok) {
      std::fclose(f);
      return STQ_ERR_INVALID_ARGUMENT;
    }
    table[key] = e;
  }
  const bool clean_end = r == EOF;
  std::fclose(f);
  if (!clean_end) return STQ_ERR_INVALID_ARGUMENT;
  for (const auto &kv : table) tune_table[kv.first] = kv.second;
  return STQ_OK;
}

}  // namespace stqamd
