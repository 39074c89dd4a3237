// Calls the CPU-entry drop-in exactly as the reference's app/main.cpp:138-153
// does (sweep_cpu_single_batch until a batch returns no pair), reading boxes
// from a binary file (N x 72-byte FP64 records, id = position) and writing
// every pair returned as int32 (a.id, b.id) pairs.
// Usage: dropin_cpu_demo boxes.bin pairs.bin [n]
#include <stq/cpu/stq_cpu_dropin.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace stq::cpu;

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  std::FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  std::vector<Aabb> boxes;
  stq_aabb_f64 r;
  while (std::fread(&r, sizeof(r), 1, f) == 1) {
    Aabb b((int)boxes.size(), {r.vertex_ids[0], r.vertex_ids[1], r.vertex_ids[2]},
           {r.min[0], r.min[1], r.min[2]}, {r.max[0], r.max[1], r.max[2]});
    boxes.push_back(b);
  }
  std::fclose(f);
  const int N = (int)boxes.size();
  int n = argc > 3 ? std::atoi(argv[3]) : N;
  std::vector<Aabb> boxes_batching(boxes);
  std::vector<std::pair<int, int>> overlaps, all;
  std::size_t count = 0;
  int batches = 1;
  sweep_cpu_single_batch(boxes_batching, n, N, overlaps);
  while (overlaps.size()) {
    count += overlaps.size();
    all.insert(all.end(), overlaps.begin(), overlaps.end());
    sweep_cpu_single_batch(boxes_batching, n, N, overlaps);
    ++batches;
  }
  std::printf("boxes %d final count %zu batches %d left %zu\n", N, count, batches,
              boxes_batching.size());
  std::FILE *o = std::fopen(argv[2], "wb");
  for (auto &p : all) {
    int v[2] = {p.first, p.second};
    std::fwrite(v, sizeof(int), 2, o);
  }
  std::fclose(o);
  return 0;
}
