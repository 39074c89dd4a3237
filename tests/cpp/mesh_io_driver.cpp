// Runs the host-side readers of libstq_amd (csrc/mesh_io.cpp, compiled into
// this program) on every file named on the command line: .obj and .ply
// through stq_mesh_read (and stq_mesh_edges on what it read), .json through
// stq_recall with a few pairs.  Prints one line per file: the path, the
// status and the sizes.  Built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitizers.py.
#include <stq.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

int main(int argc, char **argv) {
  for (int a = 1; a < argc; ++a) {
    const std::string p(argv[a]);
    const std::string ext = p.size() > 4 ? p.substr(p.size() - 4) : "";
    if (ext == "json") {
      const stq_pair pairs[3] = {{0, 1}, {1, 2}, {5, 3}};
      int64_t tp = -1, tot = -1;
      const int rc = stq_recall(pairs, 3, p.c_str(), 1, &tp, &tot);
      std::printf("%s %d %lld %lld\n", p.c_str(), rc, (long long)tp, (long long)tot);
      continue;
    }
    double *v = nullptr;
    int32_t *f = nullptr;
    int64_t nv = -1, nf = -1, ne = -1;
    const int rc = stq_mesh_read(p.c_str(), &v, &nv, &f, &nf);
    if (rc == STQ_OK) {
      int32_t *e = nullptr;
      const int rce = stq_mesh_edges(f, nf, &e, &ne);
      if (rce != STQ_OK) ne = -rce;
      // every index is read back (a bad one would be seen by the sanitizers)
      long long sum = 0;
      for (int64_t i = 0; i < 3 * nf; ++i) sum += f[i];
      for (int64_t i = 0; i < 2 * ne; ++i) sum += e[i];
      stq_free_host(e);
      if (sum < 0) std::printf("negative index sum\n");
    }
    std::printf("%s %d %lld %lld %lld\n", p.c_str(), rc, (long long)nv, (long long)nf, (long long)ne);
    stq_free_host(v);
    stq_free_host(f);
  }
  return 0;
}

// stq_free_host lives in api.cpp with the GPU code; the readers' buffers come
// from malloc.
extern "C" void stq_free_host(void *h_ptr) { std::free(h_ptr); }
