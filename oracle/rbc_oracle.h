/* CPU oracle for the RBC engine -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may call
 * these functions, and only as the checker or the CPU baseline.  The product
 * library (librbc_hip.so) never links or calls them.
 *
 * Every function restates a piece of the reference (lcayton/RandomBallCover-GPU
 * v0.2.6) in plain C, citing the file:line it follows, or is an FP64 ground truth.
 * Conventions shared with the product:
 *   - matrices are row-major with leading dimension ld (floats);
 *   - metric 0 = L2 (squared Euclidean, defs.h:21), 1 = L1 (sum |.|, defs.h:20);
 *   - k-NN results are ordered by (distance, index) ascending, i.e. ties are
 *     broken by the lowest index; unfilled slots hold (FLT_MAX, UINT32_MAX).
 */
#ifndef RBC_ORACLE_H
#define RBC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FP32 distance as the reference's device kernels compute it: one fused
 * multiply-add per coordinate, coordinates 0..d-1 in order
 * (kernels.cu:57-58, 193-194, 260-261, 309-310 with nvcc's default FMA contraction). */
float orc_dist_f32_fma(const float *a, const float *b, uint32_t d, int metric);
/* FP32 distance as the reference's host code computes it (utils.cu:183-191):
 * multiply, round, then add; coordinates in order. */
float orc_dist_f32_host(const float *a, const float *b, uint32_t d, int metric);
/* FP64 ground truth: t = (double)a - (double)b, acc += t*t (or |t|), in order,
 * no contraction. */
double orc_dist_f64(const float *a, const float *b, uint32_t d, int metric);
/* FP32 pairwise (tree) summation of the per-coordinate terms t_c = fl(fl(a_c-b_c)^2)
 * (or |a_c-b_c|): the order the product's exact re-rank uses.  With P the next power
 * of two >= max(d,64) and t_c = 0 for c >= d, for s = P/2, P/4, ..., 1:
 * t[i] = t[i] + t[i+s] (i < s); the result is t[0]. */
float orc_dist_f32_tree(const float *a, const float *b, uint32_t d, int metric);

/* bruteK restated (brute.cu:65-88, knnKernel kernels.cu:229-282): top-k of each
 * query over all n points with FMA-order FP32 distances. */
void orc_brute_k_f32(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, uint32_t k, uint32_t *ids,
                     float *dists, uint32_t ldo, int threads);
/* FP64 ground truth k-NN (ties by lowest index). dists in double. */
void orc_brute_k_f64(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, uint32_t k, uint32_t *ids,
                     double *dists, uint32_t ldo, int threads);
/* bruteCPU restated (brute.cu:91-112): 1-NN with host-order distances and a strict
 * '<' scan, so the lowest index wins ties. */
void orc_brute_cpu(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                   uint32_t ldq, uint32_t d, int metric, uint32_t *ids, int threads);
/* 1-NN with FMA-order FP32 distances, lowest index on ties (nnKernel role,
 * kernels.cu:165-222, under the lowest-index contract). */
void orc_nn_f32(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                uint32_t ldq, uint32_t d, int metric, uint32_t *ids, float *dists, int threads);
/* rangeCountKernel restated (kernels.cu:426-471): #points with dist < range. */
void orc_range_count(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, const float *ranges,
                     uint32_t *counts, int threads);

/* One-shot build restated (rbc.cu:126-208): D = rep x point distances
 * (dist1Kernel kernels.cu:285-317), per-rep bisection range (findRangeKernel
 * kernels.cu:322-414, including the 4-row block coupling of its continue flags),
 * flags D < range (kernels.cu:417-423), exclusive scan + scatter (sKernel.cu:14-153)
 * giving ascending-index lists.  reps are the rows x[rep_ids[i]].
 * Outputs: map (R x ldmap, ldmap >= PAD(s)), counts[R], ranges[R].
 * Returns the number of rows whose bisection hit max_iters (0 = all converged). */
int orc_oneshot_build(const float *x, uint32_t n, uint32_t ldx, uint32_t d, int metric,
                      const uint32_t *rep_ids, uint32_t R, uint32_t s, uint32_t *map,
                      uint32_t ldmap, uint32_t *counts, float *ranges, uint32_t max_iters,
                      int threads);
/* kqueryRBC restated (rbc.cu:72-123 -> planKNNKernel kernels.cu:93-161): nearest rep
 * by FMA-order FP32 (lowest rep index on ties), then top-k over that rep's list. */
void orc_oneshot_kquery(const float *x, uint32_t ldx, uint32_t d, int metric,
                        const uint32_t *rep_ids, uint32_t R, const uint32_t *map,
                        uint32_t ldmap, const uint32_t *counts, const float *q, uint32_t m,
                        uint32_t ldq, uint32_t k, uint32_t *ids, float *dists, uint32_t ldo,
                        int threads);
/* orc_oneshot_kquery under a general computation plan (compPlan built from a charMatrix,
 * rbc.cu:375-431): the query's group g (nearest rep) scans the list of every rep xg with
 * plan[g * ldp + xg] != 0; the k best of the union of those lists as a multiset (a point
 * in two planned lists is inserted twice, as planKNNKernel does). */
void orc_oneshot_kquery_plan(const float *x, uint32_t ldx, uint32_t d, int metric,
                             const uint32_t *rep_ids, uint32_t R, const uint32_t *map,
                             uint32_t ldmap, const uint32_t *counts, const uint8_t *plan,
                             uint32_t ldp, const float *q, uint32_t m, uint32_t ldq, uint32_t k,
                             uint32_t *ids, float *dists, uint32_t ldo, int threads);

/* Exact-RBC build pieces: owner[i] = nearest rep of x_i (FMA-order FP32, lowest rep
 * index on ties); radius[j] = max over owned points of the FP64 metric distance
 * (sqrt of the squared sum for L2), rounded up to float. */
void orc_exact_owners(const float *x, uint32_t n, uint32_t ldx, uint32_t d, int metric,
                      const uint32_t *rep_ids, uint32_t R, uint32_t *owner, int threads);
void orc_exact_radii(const float *x, uint32_t n, uint32_t ldx, uint32_t d, int metric,
                     const uint32_t *rep_ids, uint32_t R, const uint32_t *owner,
                     float *radius);

/* CPU exact RBC k-NN (triangle and Voronoi-plane pruning, AVX2 over points, OpenMP over
 * queries, L2): the results of orc_brute_k_f32 from the lists of an exact index (perm,
 * off[R+1], reps, radii, FP64 rep-to-rep distances rr from orc_rep_dist2) with the points
 * in orc_rbc_soa's layout (orc_rbc_soa_size floats). */
size_t orc_rbc_soa_size(const uint32_t *off, uint32_t R, uint32_t d);
void orc_rbc_soa(const float *x, uint32_t ldx, uint32_t d, const uint32_t *perm,
                 const uint32_t *off, uint32_t R, float *soa);
void orc_rep_dist2(const float *reps, uint32_t ldr, uint32_t R, uint32_t d, double *rr, int threads);
void orc_rbc_exact_knn(const float *soa, const float *q, uint32_t m, uint32_t ldq, uint32_t d,
                       const float *reps, uint32_t ldr, uint32_t R, const uint32_t *perm,
                       const uint32_t *off, const float *radii, const double *rr, uint32_t k,
                       uint32_t *ids, float *dists, uint32_t ldo, int threads, double *pairs);

#ifdef __cplusplus
}
#endif
#endif
