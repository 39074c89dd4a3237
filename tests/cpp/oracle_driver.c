/* Exercises the CPU oracle (oracle/stq_oracle.c, compiled into this program)
 * on random scenes and edge cases: the sweep on every axis and row count
 * against the brute force, canonicalisation, the swept-box build with good
 * and bad indices, the orientation.  Exit status 0 when every result agrees.
 * Built with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitizers.py. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t id, pad0;
  double min[3], max[3];
  int32_t vids[3], ref_id;
} box_t;

int64_t oracle_sweep_rows(const box_t *, int64_t, int, int64_t, int, int32_t **, double *);
int64_t oracle_bruteforce(const box_t *, int64_t, int, int32_t **);
int64_t oracle_canonicalize(int32_t *, int64_t);
int oracle_build_swept_boxes(const double *, const double *, int64_t, const int32_t *, int64_t,
                             const int32_t *, int64_t, double, int, box_t *);
int oracle_orient_pairs(const box_t *, int64_t, int32_t *, int64_t);
void oracle_free(void *);

static uint64_t rng = 0x9e3779b97f4a7c15ull;
static double urand(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

int main(void) {
  int bad = 0;
  for (int trial = 0; trial < 40; ++trial) {
    const int64_t n = trial < 3 ? trial : 50 + (int64_t)(urand() * 900);
    box_t *b = (box_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(box_t));
    for (int64_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        const double c = urand(), h = 0.02 * urand();
        b[i].min[k] = trial % 5 == 0 ? 0.5 : c - h; /* some scenes of equal minima */
        b[i].max[k] = c + h > b[i].min[k] ? c + h : b[i].min[k];
      }
      const int kind = (int)(urand() * 3);
      const int32_t v = (int32_t)(urand() * 400);
      b[i].vids[0] = v;
      b[i].vids[1] = kind == 0 ? -v - 1 : v + 1;
      b[i].vids[2] = kind == 2 ? v + 2 : (kind == 1 ? -v - 1 : -v - 1);
    }
    int32_t *ref = NULL;
    const int64_t nb = oracle_bruteforce(b, n, 2, &ref);
    const int64_t nrc = oracle_canonicalize(ref, nb);
    for (int axis = 0; axis < 3; ++axis) {
      int32_t *got = NULL;
      double st = 0;
      const int64_t ng = oracle_sweep_rows(b, n, axis, n, 1 + trial % 3, &got, &st);
      const int64_t cg = oracle_canonicalize(got, ng);
      if (cg != nrc || (nrc && memcmp(got, ref, (size_t)nrc * 8) != 0)) {
        printf("sweep mismatch: trial %d axis %d (%lld vs %lld)\n", trial, axis, (long long)cg,
               (long long)nrc);
        bad = 1;
      }
      oracle_free(got);
      /* a partial row range never reads outside the boxes */
      got = NULL;
      const int64_t np = oracle_sweep_rows(b, n, axis, n / 3, 2, &got, &st);
      if (np > ng) bad = 1;
      oracle_free(got);
    }
    oracle_free(ref);
    free(b);
  }
  /* swept boxes of a tiny mesh, then a bad index */
  const double v0[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0}, v1[9] = {0, 0, 1, 1, 0, 1, 0, 1, 1};
  const int32_t e[6] = {0, 1, 1, 2, 2, 0}, f[3] = {0, 1, 2}, fbad[3] = {0, 1, 7};
  box_t out[7];
  if (oracle_build_swept_boxes(v0, v1, 3, e, 3, f, 1, 0.01, 1, out) != 0) bad = 1;
  if (oracle_build_swept_boxes(v0, v1, 3, e, 3, fbad, 1, 0.0, 0, out) == 0) bad = 1;
  int32_t pr[4] = {6, 0, 3, 4};
  if (oracle_orient_pairs(out, 7, pr, 2) != 0) bad = 1;
  int32_t prb[2] = {0, 99};
  if (oracle_orient_pairs(out, 7, prb, 1) == 0) bad = 1;
  printf(bad ? "oracle driver: FAILED\n" : "oracle driver: ok\n");
  return bad;
}
