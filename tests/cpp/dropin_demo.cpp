// Calls the C++ drop-in as the reference's app/main.cu:141-144 calls
// run_sweep_sharedqueue, in a loop over tidstart (the loop a caller needs
// once the cutoff shrinks, src/simulation.cu:747), reading boxes from a
// binary file (N x 72-byte FP64 Aabb records) and writing every pair returned
// as int32 pairs.
// Usage: dropin_demo boxes.bin pairs.bin [memlimit_gb [cutoff]]
#define KEEP_CPU_OVERLAPS
#include <stq/gpu/stq_dropin.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace stq::gpu;

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  std::FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  std::vector<Aabb> boxes;
  Aabb b;
  while (std::fread(&b, sizeof(Aabb), 1, f) == 1) boxes.push_back(b);
  std::fclose(f);
  const int memlimit = argc > 3 ? std::atoi(argv[3]) : 0;
  MemHandler *memhandle = new MemHandler();
  if (argc > 4) memhandle->MAX_OVERLAP_CUTOFF = (size_t)std::atoll(argv[4]);
  std::vector<std::pair<int, int>> overlaps;
  int parallel = 0, tidstart = 0, devcount = 1, nbox = 0, calls = 0;
  const int N = (int)boxes.size();
  long long realcount = 0;
  size_t min_cutoff = (size_t)-1;
  do {
    int2 *d_overlaps = nullptr;
    int *d_count = nullptr;
    run_sweep_sharedqueue(boxes.data(), memhandle, N, nbox, overlaps, d_overlaps, d_count, parallel,
                          tidstart, devcount, memlimit);
    realcount += memhandle->realcount;
    if (memhandle->MAX_OVERLAP_CUTOFF < min_cutoff) min_cutoff = memhandle->MAX_OVERLAP_CUTOFF;
    ++calls;
    free_device(d_overlaps);
    free_device(d_count);
  } while (tidstart < N);
  std::printf("boxes %d pairs %zu realcount %lld calls %d tidstart %d threads %d cutoff %zu "
              "max_overlap_size %zu\n",
              N, overlaps.size(), realcount, calls, tidstart, parallel, min_cutoff,
              memhandle->MAX_OVERLAP_SIZE);
  std::FILE *o = std::fopen(argv[2], "wb");
  for (auto &p : overlaps) {
    int v[2] = {p.first, p.second};
    std::fwrite(v, sizeof(int), 2, o);
  }
  std::fclose(o);
  delete memhandle;
  return 0;
}
