/* CPU oracle for the RBC engine -- TEST INFRASTRUCTURE ONLY (see rbc_oracle.h).
 *
 * Built with -ffp-contract=off so that every multiply and add rounds exactly where
 * the source says; the fused multiply-adds the reference's device code gets from
 * nvcc are written out as fmaf().
 */
#include "rbc_oracle.h"

#include <float.h>
#include <immintrin.h>
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

/* Floats of the list-major, 8-point-blocked copy of the points (orc_rbc_soa). */
size_t orc_rbc_soa_size(const uint32_t *off, uint32_t R, uint32_t d) {
  size_t blocks = 0;
  for (uint32_t r = 0; r < R; ++r) blocks += (off[r + 1] - off[r] + 7) / 8;
  return blocks * 8 * d;
}

/* The CPU search's layout of the points: list by list (perm order), blocks of 8 points
 * stored coordinate-major ([d][8]), the last block of a list zero-padded. */
void orc_rbc_soa(const float *x, uint32_t ldx, uint32_t d, const uint32_t *perm,
                 const uint32_t *off, uint32_t R, float *soa) {
  size_t b = 0;
  for (uint32_t r = 0; r < R; ++r) {
    for (uint32_t p0 = off[r]; p0 < off[r + 1]; p0 += 8, ++b) {
      float *blk = soa + b * 8 * d;
      for (uint32_t l = 0; l < 8; ++l) {
        const uint32_t p = p0 + l;
        for (uint32_t c = 0; c < d; ++c)
          blk[c * 8 + l] = p < off[r + 1] ? x[(size_t)perm[p] * ldx + c] : 0.0f;
      }
    }
  }
}

/* Sequential-FMA squared distances (orc_dist_f32_fma's arithmetic, lane by lane) of
 * nb <= 4 consecutive 8-point blocks to q: four independent FMA chains per step. */
static void dist_blocks(const float *blk, uint32_t nb, const float *q, uint32_t d, float *out) {
  __m256 a0 = _mm256_setzero_ps(), a1 = a0, a2 = a0, a3 = a0;
  const size_t s = (size_t)8 * d;
  for (uint32_t c = 0; c < d; ++c) {
    const __m256 qc = _mm256_set1_ps(q[c]);
    __m256 t = _mm256_sub_ps(_mm256_loadu_ps(blk + c * 8), qc);
    a0 = _mm256_fmadd_ps(t, t, a0);
    if (nb > 1) {
      t = _mm256_sub_ps(_mm256_loadu_ps(blk + s + c * 8), qc);
      a1 = _mm256_fmadd_ps(t, t, a1);
    }
    if (nb > 2) {
      t = _mm256_sub_ps(_mm256_loadu_ps(blk + 2 * s + c * 8), qc);
      a2 = _mm256_fmadd_ps(t, t, a2);
    }
    if (nb > 3) {
      t = _mm256_sub_ps(_mm256_loadu_ps(blk + 3 * s + c * 8), qc);
      a3 = _mm256_fmadd_ps(t, t, a3);
    }
  }
  _mm256_storeu_ps(out, a0);
  _mm256_storeu_ps(out + 8, a1);
  _mm256_storeu_ps(out + 16, a2);
  _mm256_storeu_ps(out + 24, a3);
}

typedef struct {
  float v;
  uint32_t r;
} orc_key;

static int key_cmp(const void *a, const void *b) {
  const orc_key *x = (const orc_key *)a, *y = (const orc_key *)b;
  if (x->v != y->v) return x->v < y->v ? -1 : 1;
  return x->r < y->r ? -1 : (x->r > y->r);
}

/* Squared FP64 distances between the R representatives (rr[R][R]), for the plane bound. */
void orc_rep_dist2(const float *reps, uint32_t ldr, uint32_t R, uint32_t d, double *rr, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 8)
  for (long i = 0; i < (long)R; ++i)
    for (uint32_t j = 0; j < R; ++j)
      rr[(size_t)i * R + j] = orc_dist_f64(reps + (size_t)i * ldr, reps + (size_t)j * ldr, d, 0);
}

/* CPU exact RBC k-NN: the reference's RBC search structure (rbc.cu:19-123, representatives
 * and their lists) restated on the host as an exact search with the triangle-inequality
 * pruning of the exact RBC (computeRadii, rbc.cu:250-256; deprecated.cu:354-464) and the
 * Voronoi-plane bound of an exact (nearest-rep) assignment, vectorized over points (AVX2,
 * 8-point blocks of orc_rbc_soa's layout) and parallel over queries (OpenMP).  Each query
 * computes its distances to the R representatives, visits the lists in ascending
 * representative distance and skips list j when every point of it is provably farther than
 * its current k-th result.  With e = (d' + 4) 2^-24 (1.01), d' = d rounded up to a multiple
 * of 4 (the relative error of an FP32 sequential-FMA distance, and of the FP32 owner test
 * that made the lists), the bounds on the true distance of q to a point of list j are
 *   triangle: sqrt(Dqr_j (1 - e)) - psi_j,
 *   plane:    (Dqr_j (1 - e) - Dqr_i (1 + e) - 2 e psi_j^2) / (2 |r_j - r_i|)
 * for the query's P nearest reps r_i (P = 4; a point x owned by r_j has |x - r_j|^2 <=
 * |x - r_i|^2 + 2 e psi_j^2 for every rep), and the list is skipped when the largest bound
 * L > 0 has L^2 (1 - e) above the k-th FP32 distance.  Inside a list it keeps the k best
 * (FP32 sequential-FMA distances, lowest original index among equals), so the results
 * equal orc_brute_k_f32's.  L2 only.  perm maps permuted rows to original ids, off[R+1]
 * are the list offsets, reps[R][ldr], radii[R] the list radii (max true distance of a
 * list's points to its rep, rounded up), rr the FP64 rep-to-rep squared distances
 * (orc_rep_dist2; NULL: triangle bound only), soa the points in orc_rbc_soa's layout.
 * *pairs receives the number of (query, point) distances evaluated. */
void orc_rbc_exact_knn(const float *soa, const float *q, uint32_t m, uint32_t ldq, uint32_t d,
                       const float *reps, uint32_t ldr, uint32_t R, const uint32_t *perm,
                       const uint32_t *off, const float *radii, const double *rr, uint32_t k,
                       uint32_t *ids, float *dists, uint32_t ldo, int threads, double *pairs) {
  set_threads(threads);
  const double eps = ((double)((d + 3u) & ~3u) + 4.0) * 0x1p-24 * 1.01;
  const uint32_t P = R < 4 ? R : 4;
  size_t *boff = (size_t *)malloc(sizeof(size_t) * (R + 1));
  boff[0] = 0;
  for (uint32_t r = 0; r < R; ++r) boff[r + 1] = boff[r] + (off[r + 1] - off[r] + 7) / 8;
  double total = 0.0;
#pragma omp parallel reduction(+ : total)
  {
    orc_key *key = (orc_key *)malloc(sizeof(orc_key) * R);
    float dd[32];
#pragma omp for schedule(dynamic, 1)
    for (long i = 0; i < (long)m; ++i) {
      const float *qi = q + (size_t)i * ldq;
      float *dv = dists + (size_t)i * ldo;
      uint32_t *iv = ids + (size_t)i * ldo;
      for (uint32_t j = 0; j < k; ++j) {
        dv[j] = FLT_MAX;
        iv[j] = ORC_DUMMY;
      }
      for (uint32_t r = 0; r < R; ++r) {
        key[r].v = orc_dist_f32_fma(reps + (size_t)r * ldr, qi, d, 0);
        key[r].r = r;
      }
      qsort(key, R, sizeof(orc_key), key_cmp);
      for (uint32_t a = 0; a < R; ++a) {
        const uint32_t r = key[a].r;
        if (dv[k - 1] < FLT_MAX) {
          const double psi = (double)radii[r], Dlo = (double)key[a].v * (1.0 - eps);
          double L = sqrt(Dlo) - psi;
          for (uint32_t pi = 0; rr && pi < P; ++pi) {
            const uint32_t ri = key[pi].r;
            const double rij = rr[(size_t)ri * R + r];
            if (ri == r || !(rij > 0.0)) continue;
            const double num = Dlo - (double)key[pi].v * (1.0 + eps) - 2.0 * eps * psi * psi;
            const double pl = num / (2.0 * sqrt(rij) * (1.0 + 1e-12));
            if (pl > L) L = pl;
          }
          L *= 1.0 - 1e-12;
          if (L > 0.0 && L * L * (1.0 - eps) * (1.0 - 1e-12) > (double)dv[k - 1]) continue;
        }
        const uint32_t len = off[r + 1] - off[r];
        const uint32_t nblk = (len + 7) / 8;
        for (uint32_t b = 0; b < nblk; b += 4) {
          const uint32_t nb = nblk - b < 4 ? nblk - b : 4;
          dist_blocks(soa + (boff[r] + b) * 8 * d, nb, qi, d, dd);
          const uint32_t p0 = b * 8, cnt = len - p0 < 32 ? len - p0 : 32;
          for (uint32_t l = 0; l < cnt; ++l)
            if (dd[l] <= dv[k - 1]) topk_insert_f(dv, iv, k, dd[l], perm[off[r] + p0 + l]);
        }
        total += (double)len;
      }
    }
    free(key);
  }
  free(boff);
  if (pairs) *pairs = total;
}
