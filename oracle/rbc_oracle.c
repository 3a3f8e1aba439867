/* CPU oracle for the RBC engine -- TEST INFRASTRUCTURE ONLY (see rbc_oracle.h).
 *
 * Built with -ffp-contract=off so that every multiply and add rounds exactly where
 * the source says; the fused multiply-adds the reference's device code gets from
 * nvcc are written out as fmaf().
 */
#include "rbc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_DUMMY UINT32_MAX

static void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

/* kernels.cu:57-58 etc.: ans += DIST(x, q) with DIST = (x-q)*(x-q) (defs.h:21),
 * contracted to one FMA per coordinate; L1 is fabs(x-q) (defs.h:20). */
float orc_dist_f32_fma(const float *a, const float *b, uint32_t d, int metric) {
  float acc = 0.0f;
  if (metric == 0) {
    for (uint32_t c = 0; c < d; ++c) {
      float t = a[c] - b[c];
      acc = fmaf(t, t, acc);
    }
  } else {
    for (uint32_t c = 0; c < d; ++c) acc = acc + fabsf(a[c] - b[c]);
  }
  return acc;
}

/* utils.cu:183-191 (distVec): host code, product rounded before the add. */
float orc_dist_f32_host(const float *a, const float *b, uint32_t d, int metric) {
  float acc = 0.0f;
  for (uint32_t c = 0; c < d; ++c) {
    float t = a[c] - b[c];
    float term = metric == 0 ? t * t : fabsf(t);
    acc = acc + term;
  }
  return acc;
}

double orc_dist_f64(const float *a, const float *b, uint32_t d, int metric) {
  double acc = 0.0;
  for (uint32_t c = 0; c < d; ++c) {
    double t = (double)a[c] - (double)b[c];
    double term = metric == 0 ? t * t : fabs(t);
    acc = acc + term;
  }
  return acc;
}

float orc_dist_f32_tree(const float *a, const float *b, uint32_t d, int metric) {
  uint32_t P = 64;
  while (P < d) P <<= 1;
  float t[4096];
  if (P > 4096) return NAN;
  for (uint32_t c = 0; c < P; ++c) {
    if (c < d) {
      float u = a[c] - b[c];
      t[c] = metric == 0 ? u * u : fabsf(u);
    } else {
      t[c] = 0.0f;
    }
  }
  for (uint32_t s = P / 2; s >= 1; s >>= 1)
    for (uint32_t i = 0; i < s; ++i) t[i] = t[i] + t[i + s];
  return t[0];
}

/* Insert (v, id) into the ascending (dist, index) list of length k if it belongs. */
static void topk_insert_f(float *dv, uint32_t *iv, uint32_t k, float v, uint32_t id) {
  if (v > dv[k - 1] || (v == dv[k - 1] && id >= iv[k - 1])) return;
  int p = (int)k - 1;
  while (p > 0 && (dv[p - 1] > v || (dv[p - 1] == v && iv[p - 1] > id))) {
    dv[p] = dv[p - 1];
    iv[p] = iv[p - 1];
    --p;
  }
  dv[p] = v;
  iv[p] = id;
}

static void topk_insert_d(double *dv, uint32_t *iv, uint32_t k, double v, uint32_t id) {
  if (v > dv[k - 1] || (v == dv[k - 1] && id >= iv[k - 1])) return;
  int p = (int)k - 1;
  while (p > 0 && (dv[p - 1] > v || (dv[p - 1] == v && iv[p - 1] > id))) {
    dv[p] = dv[p - 1];
    iv[p] = iv[p - 1];
    --p;
  }
  dv[p] = v;
  iv[p] = id;
}

void orc_brute_k_f32(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, uint32_t k, uint32_t *ids,
                     float *dists, uint32_t ldo, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < (long)m; ++i) {
    float *dv = dists + (size_t)i * ldo;
    uint32_t *iv = ids + (size_t)i * ldo;
    for (uint32_t j = 0; j < k; ++j) {
      dv[j] = FLT_MAX;
      iv[j] = ORC_DUMMY;
    }
    const float *qi = q + (size_t)i * ldq;
    for (uint32_t j = 0; j < n; ++j)
      topk_insert_f(dv, iv, k, orc_dist_f32_fma(x + (size_t)j * ldx, qi, d, metric), j);
  }
}

void orc_brute_k_f64(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, uint32_t k, uint32_t *ids,
                     double *dists, uint32_t ldo, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
  for (long i = 0; i < (long)m; ++i) {
    double *dv = dists + (size_t)i * ldo;
    uint32_t *iv = ids + (size_t)i * ldo;
    for (uint32_t j = 0; j < k; ++j) {
      dv[j] = DBL_MAX;
      iv[j] = ORC_DUMMY;
    }
    const float *qi = q + (size_t)i * ldq;
    for (uint32_t j = 0; j < n; ++j)
      topk_insert_d(dv, iv, k, orc_dist_f64(x + (size_t)j * ldx, qi, d, metric), j);
  }
}

void orc_brute_cpu(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                   uint32_t ldq, uint32_t d, int metric, uint32_t *ids, int threads) {
  set_threads(threads);
  /* brute.cu:99-108: dtoNN starts at MAX_REAL, NN at 0, strict '<' update. */
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < (long)m; ++i) {
    float best = FLT_MAX;
    uint32_t arg = 0;
    for (uint32_t j = 0; j < n; ++j) {
      float t = orc_dist_f32_host(q + (size_t)i * ldq, x + (size_t)j * ldx, d, metric);
      if (t < best) {
        best = t;
        arg = j;
      }
    }
    ids[i] = arg;
  }
}

void orc_nn_f32(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                uint32_t ldq, uint32_t d, int metric, uint32_t *ids, float *dists, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < (long)m; ++i) {
    float best = FLT_MAX;
    uint32_t arg = ORC_DUMMY;
    for (uint32_t j = 0; j < n; ++j) {
      float t = orc_dist_f32_fma(x + (size_t)j * ldx, q + (size_t)i * ldq, d, metric);
      if (t < best) {
        best = t;
        arg = j;
      }
    }
    ids[i] = arg;
    if (dists) dists[i] = best;
  }
}

void orc_range_count(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, const float *ranges,
                     uint32_t *counts, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < (long)m; ++i) {
    uint32_t cnt = 0;
    for (uint32_t j = 0; j < n; ++j)
      cnt += orc_dist_f32_fma(x + (size_t)j * ldx, q + (size_t)i * ldq, d, metric) < ranges[i];
    counts[i] = cnt;
  }
}

/* findRangeKernel (kernels.cu:322-414) for one block of four consecutive rows
 * r0..r0+3 of D (rows >= R are the padding rows of the reference's chunk). */
static int find_range_block(const float *D, uint32_t n, uint32_t R, uint32_t r0,
                            uint32_t cntWant, float *ranges, uint32_t max_iters) {
  const uint32_t LB = (90 * cntWant) / 100; /* kernels.cu:329 */
  const uint32_t UB = cntWant;              /* kernels.cu:330 */
  float smin[4], smax[4], rg[4] = {0, 0, 0, 0};
  int cont[4];
  for (int ro = 0; ro < 4; ++ro) {
    uint32_t row = r0 + ro;
    float mn = FLT_MAX, mx = 0.0f; /* kernels.cu:335-336 */
    if (row < R) {
      const float *Dr = D + (size_t)row * n;
      for (uint32_t c = 0; c < n; ++c) {
        float t = Dr[c];
        mn = (t <= mn) ? t : mn; /* MIN, defs.h:48 */
        mx = (t > mx) ? t : mx;  /* MAX, defs.h:47 */
      }
    }
    smin[ro] = mn;
    smax[ro] = mx;
    cont[ro] = 1; /* kernels.cu:365-366 */
  }
  uint32_t it = 0;
  do {
    ++it;
    for (int ro = 0; ro < 4; ++ro) {
      uint32_t row = r0 + ro;
      if (cont[ro]) rg[ro] = (smax[ro] + smin[ro]) / 2.0f; /* kernels.cu:372-373 */
      uint32_t cc = 0;
      if (row < R) {
        const float *Dr = D + (size_t)row * n;
        for (uint32_t c = 0; c < n; ++c) cc += Dr[c] <= rg[ro]; /* kernels.cu:377 */
      }
      if (cc < cntWant) /* kernels.cu:390-395 */
        smin[ro] = rg[ro];
      else
        smax[ro] = rg[ro];
      cont[ro] = row < R && (cc < LB || cc > UB); /* kernels.cu:398-399 */
    }
    /* kernels.cu:403-407: the OR-reduction writes into rows 0 and 1, which then
     * decide whether those rows recompute their midpoint next iteration. */
    cont[0] |= cont[2];
    cont[1] |= cont[3];
    cont[0] |= cont[1];
  } while (cont[0] && it < max_iters);
  for (int ro = 0; ro < 4; ++ro)
    if (r0 + ro < R) ranges[r0 + ro] = rg[ro];
  return cont[0] ? 1 : 0;
}

int orc_oneshot_build(const float *x, uint32_t n, uint32_t ldx, uint32_t d, int metric,
                      const uint32_t *rep_ids, uint32_t R, uint32_t s, uint32_t *map,
                      uint32_t ldmap, uint32_t *counts, float *ranges, uint32_t max_iters,
                      int threads) {
  set_threads(threads);
  float *D = (float *)malloc((size_t)R * n * sizeof(float));
  if (!D) return -1;
  /* dist1Kernel (kernels.cu:285-317): D[r][j] = sum_c DIST(rep_c, x_c). */
#pragma omp parallel for schedule(dynamic, 1)
  for (long r = 0; r < (long)R; ++r) {
    const float *rp = x + (size_t)rep_ids[r] * ldx;
    for (uint32_t j = 0; j < n; ++j)
      D[(size_t)r * n + j] = orc_dist_f32_fma(rp, x + (size_t)j * ldx, d, metric);
  }
  int bad = 0;
  uint32_t nblk = (R + 3) / 4;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (long b = 0; b < (long)nblk; ++b)
    bad += find_range_block(D, n, R, (uint32_t)b * 4, s, ranges, max_iters);
  /* rangeSearchKernel + sumKernel + buildMapKernel + getCountsKernel: the list of
   * rep r is the ascending set {j < n : D[r][j] < range[r]} (kernels.cu:421,
   * sKernel.cu:144-153); its length is the count (sKernel.cu:136-141). */
  for (uint32_t r = 0; r < R; ++r) {
    uint32_t c = 0;
    for (uint32_t j = 0; j < n; ++j) {
      if (D[(size_t)r * n + j] < ranges[r]) {
        if (c < ldmap) map[(size_t)r * ldmap + c] = j;
        ++c;
      }
    }
    for (uint32_t t = c; t < ldmap; ++t) map[(size_t)r * ldmap + t] = 0;
    counts[r] = c;
  }
  free(D);
  return bad;
}

void orc_oneshot_kquery(const float *x, uint32_t ldx, uint32_t d, int metric,
                        const uint32_t *rep_ids, uint32_t R, const uint32_t *map,
                        uint32_t ldmap, const uint32_t *counts, const float *q, uint32_t m,
                        uint32_t ldq, uint32_t k, uint32_t *ids, float *dists, uint32_t ldo,
                        int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < (long)m; ++i) {
    const float *qi = q + (size_t)i * ldq;
    /* computeReps (rbc.cu:233-247 -> nnKernel): nearest rep, lowest index on ties. */
    float best = FLT_MAX;
    uint32_t rb = 0;
    for (uint32_t r = 0; r < R; ++r) {
      float t = orc_dist_f32_fma(x + (size_t)rep_ids[r] * ldx, qi, d, metric);
      if (t < best) {
        best = t;
        rb = r;
      }
    }
    float *dv = dists + (size_t)i * ldo;
    uint32_t *iv = ids + (size_t)i * ldo;
    for (uint32_t j = 0; j < k; ++j) {
      dv[j] = FLT_MAX;
      iv[j] = ORC_DUMMY;
    }
    /* planKNNKernel (kernels.cu:122-151): scan the rep's list. */
    uint32_t cnt = counts[rb] < ldmap ? counts[rb] : ldmap;
    for (uint32_t t = 0; t < cnt; ++t) {
      uint32_t j = map[(size_t)rb * ldmap + t];
      topk_insert_f(dv, iv, k, orc_dist_f32_fma(x + (size_t)j * ldx, qi, d, metric), j);
    }
  }
}

void orc_oneshot_kquery_plan(const float *x, uint32_t ldx, uint32_t d, int metric,
                             const uint32_t *rep_ids, uint32_t R, const uint32_t *map,
                             uint32_t ldmap, const uint32_t *counts, const uint8_t *plan,
                             uint32_t ldp, const float *q, uint32_t m, uint32_t ldq, uint32_t k,
                             uint32_t *ids, float *dists, uint32_t ldo, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < (long)m; ++i) {
    const float *qi = q + (size_t)i * ldq;
    /* computeReps (rbc.cu:233-247): nearest rep, lowest index on ties; it names the
     * query's group. */
    float best = FLT_MAX;
    uint32_t g = 0;
    for (uint32_t r = 0; r < R; ++r) {
      float t = orc_dist_f32_fma(x + (size_t)rep_ids[r] * ldx, qi, d, metric);
      if (t < best) {
        best = t;
        g = r;
      }
    }
    float *dv = dists + (size_t)i * ldo;
    uint32_t *iv = ids + (size_t)i * ldo;
    for (uint32_t j = 0; j < k; ++j) {
      dv[j] = FLT_MAX;
      iv[j] = ORC_DUMMY;
    }
    /* planNNKernel / planKNNKernel (kernels.cu:38-72, 120-151): every list the plan row
     * of the group names (initCompPlan, rbc.cu:405-412, keeps them in ascending order),
     * every point merged without a check for points seen before. */
    for (uint32_t xg = 0; xg < R; ++xg) {
      if (!plan[(size_t)g * ldp + xg]) continue;
      uint32_t cnt = counts[xg] < ldmap ? counts[xg] : ldmap;
      for (uint32_t t = 0; t < cnt; ++t) {
        uint32_t j = map[(size_t)xg * ldmap + t];
        topk_insert_f(dv, iv, k, orc_dist_f32_fma(x + (size_t)j * ldx, qi, d, metric), j);
      }
    }
  }
}

void orc_exact_owners(const float *x, uint32_t n, uint32_t ldx, uint32_t d, int metric,
                      const uint32_t *rep_ids, uint32_t R, uint32_t *owner, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (long i = 0; i < (long)n; ++i) {
    float best = FLT_MAX;
    uint32_t rb = 0;
    for (uint32_t r = 0; r < R; ++r) {
      float t = orc_dist_f32_fma(x + (size_t)i * ldx, x + (size_t)rep_ids[r] * ldx, d, metric);
      if (t < best) {
        best = t;
        rb = r;
      }
    }
    owner[i] = rb;
  }
}

void orc_exact_radii(const float *x, uint32_t n, uint32_t ldx, uint32_t d, int metric,
                     const uint32_t *rep_ids, uint32_t R, const uint32_t *owner,
                     float *radius) {
  double *rad = (double *)calloc(R, sizeof(double));
  for (uint32_t i = 0; i < n; ++i) {
    double v = orc_dist_f64(x + (size_t)i * ldx, x + (size_t)rep_ids[owner[i]] * ldx, d, metric);
    if (metric == 0) v = sqrt(v);
    if (v > rad[owner[i]]) rad[owner[i]] = v;
  }
  for (uint32_t r = 0; r < R; ++r) {
    float f = (float)rad[r];
    if ((double)f < rad[r]) f = nextafterf(f, FLT_MAX);
    radius[r] = f;
  }
  free(rad);
}
