/* Runs every oracle routine on small, ragged inputs (n and d not multiples of 4 or 16,
 * k up to KMAX, empty lists) under AddressSanitizer and UndefinedBehaviorSanitizer, and
 * checks the identities that tie them together:
 *   orc_rbc_exact_knn == orc_brute_k_f32 (ids and distances, with and without planes);
 *   orc_oneshot_kquery_plan with the identity plan == orc_oneshot_kquery;
 *   orc_exact_owners == orc_nn_f32 against the rep rows;
 *   orc_brute_cpu, orc_range_count and orc_brute_k_f64 agree on their shared facts.
 * Prints "ok" and exits 0, or names the failed check and exits 1. */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rbc_oracle.h"

static int failures = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);      \
      ++failures;                                                         \
    }                                                                     \
  } while (0)

static float frand(unsigned *s) {
  *s = *s * 1664525u + 1013904223u;
  return (float)((*s >> 8) & 0xFFFF) / 65536.0f;
}

int main(void) {
  const uint32_t n = 1001, d = 13, m = 77, R = 29, k = 32, s = 90;
  unsigned seed = 12345;
  float *x = malloc(sizeof(float) * n * d), *q = malloc(sizeof(float) * m * d);
  for (uint32_t i = 0; i < n * d; ++i) x[i] = (float)(int)(frand(&seed) * 6.0f); /* many ties */
  for (uint32_t i = 0; i < m * d; ++i) q[i] = (float)(int)(frand(&seed) * 6.0f);
  uint32_t rep[29];
  for (uint32_t r = 0; r < R; ++r) rep[r] = r * 34 + 3;

  uint32_t *bi = malloc(4 * m * k), *ei = malloc(4 * m * k), *bi64 = malloc(4 * m * k);
  float *bd = malloc(4 * m * k), *ed = malloc(4 * m * k);
  double *bd64 = malloc(8 * m * k);
  orc_brute_k_f32(x, n, d, q, m, d, d, 0, k, bi, bd, k, 2);
  orc_brute_k_f64(x, n, d, q, m, d, d, 0, k, bi64, bd64, k, 2);
  for (uint32_t i = 0; i < m; ++i)
    for (uint32_t j = 1; j < k; ++j) CHECK(bd[i * k + j - 1] <= bd[i * k + j]);
  /* Integer coordinates: FP32 and FP64 sums are exact, so both agree exactly. */
  for (uint32_t i = 0; i < m * k; ++i) CHECK(bi[i] == bi64[i] && (double)bd[i] == bd64[i]);

  /* Exact RBC over the exact owners. */
  uint32_t *owner = malloc(4 * n), *nn = malloc(4 * n);
  float *rad = malloc(4 * R), *repx = malloc(4 * R * d), *nd = malloc(4 * n);
  for (uint32_t r = 0; r < R; ++r) memcpy(repx + r * d, x + rep[r] * d, 4 * d);
  orc_exact_owners(x, n, d, d, 0, rep, R, owner, 2);
  orc_nn_f32(repx, R, d, x, n, d, d, 0, nn, nd, 2);
  for (uint32_t i = 0; i < n; ++i) CHECK(owner[i] == nn[i]);
  orc_exact_radii(x, n, d, d, 0, rep, R, owner, rad);
  uint32_t *off = calloc(R + 1, 4), *perm = malloc(4 * n), *fill = calloc(R, 4);
  for (uint32_t i = 0; i < n; ++i) ++off[owner[i] + 1];
  for (uint32_t r = 0; r < R; ++r) off[r + 1] += off[r];
  for (uint32_t i = 0; i < n; ++i) perm[off[owner[i]] + fill[owner[i]]++] = i;
  float *soa = malloc(4 * orc_rbc_soa_size(off, R, d));
  orc_rbc_soa(x, d, d, perm, off, R, soa);
  double *rr = malloc(8 * R * R), pairs = 0;
  orc_rep_dist2(repx, d, R, d, rr, 2);
  for (int planes = 0; planes < 2; ++planes) {
    orc_rbc_exact_knn(soa, q, m, d, d, repx, d, R, perm, off, rad, planes ? rr : NULL, k, ei, ed, k, 2, &pairs);
    CHECK(memcmp(ei, bi, 4 * m * k) == 0 && memcmp(ed, bd, 4 * m * k) == 0);
    CHECK(pairs > 0 && pairs <= (double)n * m);
  }

  /* bruteCPU's 1-NN distance is the k = 1 brute force distance; range counts below the
   * k-th distance are < k. */
  uint32_t *c1 = malloc(4 * m), *rc = malloc(4 * m);
  orc_brute_cpu(x, n, d, q, m, d, d, 0, c1, 2);
  float *kth = malloc(4 * m);
  for (uint32_t i = 0; i < m; ++i) {
    CHECK(orc_dist_f32_host(x + c1[i] * d, q + i * d, d, 0) == bd[i * k]);
    kth[i] = bd[i * k + k - 1];
  }
  orc_range_count(x, n, d, q, m, d, d, 0, kth, rc, 2);
  for (uint32_t i = 0; i < m; ++i) CHECK(rc[i] < k);
  CHECK(orc_dist_f32_tree(x, q, d, 0) == orc_dist_f32_fma(x, q, d, 0)); /* integers: exact */
  CHECK(orc_dist_f32_host(x, q, d, 1) == (float)orc_dist_f64(x, q, d, 1));

  /* One-shot build and queries on jittered points (the range bisection needs distinct
   * distances); the identity plan equals the plain query. */
  for (uint32_t i = 0; i < n * d; ++i) x[i] += 0.01f * frand(&seed);
  const uint32_t ldmap = (s + 15) / 16 * 16;
  uint32_t *map = malloc(4 * R * ldmap), *cnt = malloc(4 * R);
  float *ranges = malloc(4 * R);
  CHECK(orc_oneshot_build(x, n, d, d, 0, rep, R, s, map, ldmap, cnt, ranges, 1000, 2) == 0);
  for (uint32_t r = 0; r < R; ++r) CHECK(cnt[r] <= ldmap);
  uint8_t *plan = calloc(R * R, 1);
  for (uint32_t r = 0; r < R; ++r) plan[r * R + r] = 1;
  orc_oneshot_kquery(x, d, d, 0, rep, R, map, ldmap, cnt, q, m, d, k, ei, ed, k, 2);
  uint32_t *pi = malloc(4 * m * k);
  float *pd = malloc(4 * m * k);
  orc_oneshot_kquery_plan(x, d, d, 0, rep, R, map, ldmap, cnt, plan, R, q, m, d, k, pi, pd, k, 2);
  CHECK(memcmp(ei, pi, 4 * m * k) == 0 && memcmp(ed, pd, 4 * m * k) == 0);

  free(x), free(q), free(bi), free(ei), free(bi64), free(bd), free(ed), free(bd64), free(owner), free(nn);
  free(rad), free(repx), free(nd), free(off), free(perm), free(fill), free(soa), free(rr), free(map), free(cnt);
  free(ranges), free(plan), free(pi), free(pd), free(c1), free(rc), free(kth);
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("ok\n");
  return 0;
}
