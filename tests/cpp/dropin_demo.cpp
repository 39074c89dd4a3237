// Calls the C++ drop-in exactly as the reference's app/main.cu:141-144 calls
// run_sweep_sharedqueue, reading boxes from a binary file (N x 72-byte FP64
// Aabb records) and writing the pairs it returns as int32 pairs.
// Usage: dropin_demo boxes.bin pairs.bin
#define KEEP_CPU_OVERLAPS
#include <stq/gpu/stq_dropin.hpp>

#include <cstdio>
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
  MemHandler *memhandle = new MemHandler();
  std::vector<std::pair<int, int>> overlaps;
  int2 *d_overlaps = nullptr;
  int *d_count = nullptr;
  int parallel = 0, tidstart = 0, devcount = 1, nbox = 0;
  const int N = (int)boxes.size();
  run_sweep_sharedqueue(boxes.data(), memhandle, N, nbox, overlaps, d_overlaps, d_count, parallel,
                        tidstart, devcount, 0);
  std::printf("boxes %d pairs %zu realcount %d tidstart %d threads %d\n", N, overlaps.size(),
              memhandle->realcount, tidstart, parallel);
  std::FILE *o = std::fopen(argv[2], "wb");
  for (auto &p : overlaps) {
    int v[2] = {p.first, p.second};
    std::fwrite(v, sizeof(int), 2, o);
  }
  std::fclose(o);
  free_device(d_overlaps);
  free_device(d_count);
  delete memhandle;
  return 0;
}
