/* librbc_hip -- C ABI of the MI355X (gfx950) Random Ball Cover k-NN engine.
 *
 * Plain pointers and sizes only.  Every function returns an int status (RBC_OK = 0);
 * on failure rbc_last_error() gives a message for the calling thread.  Matrices are
 * row-major float32 with a leading dimension (elements per row) >= d.  `mem` tells
 * whether the data pointers are host (RBC_MEM_HOST) or device (RBC_MEM_DEVICE)
 * memory; host buffers are copied in and out by the call, device buffers are used in
 * place.  Calls on one handle are not thread-safe; all work runs on the stream
 * given (NULL = the default stream) and every call returns after its work is done.
 *
 * Each entry point names the reference interface it replaces
 * (lcayton/RandomBallCover-GPU v0.2.6, file:line).
 */
#ifndef RBC_HIP_H
#define RBC_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "rbc_defs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RBC_OK 0
#define RBC_ERR_ARG 1      /* invalid argument (replaces the reference's exit(1) paths) */
#define RBC_ERR_HIP 2      /* HIP runtime error */
#define RBC_ERR_NOMEM 3    /* device allocation failed (rbc.cu:151-154 exits instead) */
#define RBC_ERR_NOCONV 4   /* range bisection hit its iteration cap (kernels.cu:368-409 has none) */
#define RBC_ERR_LIMIT 5    /* a documented size limit was exceeded */
#define RBC_ERR_CHECK 6    /* a device bounds check failed (builds with -DRBC_CHECK only) */

#define RBC_L2 0 /* squared Euclidean, defs.h:21 */
#define RBC_L1 1 /* sum of absolute differences, defs.h:20 */

#define RBC_MEM_HOST 0
#define RBC_MEM_DEVICE 1

/* Largest dimension the kernels accept. */
#define RBC_MAX_DIM 256
/* Largest k of the exact k-query. */
#define RBC_MAX_K 64

const char *rbc_last_error(void);
/* 1 in a build with device bounds checks (-DRBC_CHECK, librbc_hip_check.so), else 0.  In
 * such a build a kernel index outside its buffer is clamped to 0 instead of read or
 * written, and the entry point returns RBC_ERR_CHECK naming the source file and line.
 * No counterpart in the reference (its debugging aid is cuda-memcheck). */
int rbc_check_enabled(void);
/* Reads element `index` of a device buffer of `bound` words (holding 1000 + i) through a
 * checked index into *value: the check build returns RBC_ERR_CHECK (and *value = 1000)
 * for index >= bound; other builds reject index >= bound before any launch. */
int rbc_check_selftest(uint32_t index, uint32_t bound, uint32_t *value);
/* Library version as 10000*major + 100*minor + patch. */
int rbc_version(void);
/* Select the HIP device used by subsequent calls of this thread (driver.cu:37-41). */
int rbc_set_device(int device);
/* The distance used by the reference-compatible entry points (the reference fixes it
 * at compile time in defs.h:18-21; here it is a process-wide setting, default RBC_L2). */
int rbc_set_metric(int metric);
int rbc_get_metric(void);

/* Measured compute peaks of the current device, for roofline reports: dense BF16 MFMA
 * (v_mfma_f32_16x16x32_bf16) and FP32 FMA (v_fma_f32) throughput in TFLOP/s, each from
 * a kernel of independent register accumulation chains timed over about 20 ms.  No
 * counterpart in the reference (its benchmarks report wall time only). */
int rbc_measure_peaks(double *bf16_mfma_tflops, double *fp32_fma_tflops, void *stream);

/* ---------------------------------------------------------------- brute force -- */

/* bruteK (brute.cu:65-88 -> knnWrap kernelWrap.cu:95-111 -> knnKernel kernels.cu:229-282):
 * for each of the m queries, the k nearest of the n database points, ascending by
 * (distance, index).  Distances are FP32 with one fused multiply-add per coordinate in
 * coordinate order, as the reference's kernel computes them.  k <= KMAX; rows of the
 * outputs have leading dimension ldo >= k; slots past n hold (MAX_REAL, DUMMY_IDX). */
int rbc_brute_k(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                uint32_t ldq, uint32_t d, int metric, uint32_t k, uint32_t *nn_ids,
                float *nn_dists, uint32_t ldo, int mem, void *stream);

/* Allocation-free variant of rbc_brute_k, with no counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree and no allocating
 * routine: its work items and per-segment partial lists live in a fixed 64 MiB device
 * array that every GPU the library runs on holds until the process ends.  x, q, nn_ids and
 * nn_dists must be device pointers; ldx and ldq multiples of 4.  Results equal
 * rbc_brute_k's.  Sizes whose partial lists cannot fit even at one segment (m * k above
 * about 8 million) are rejected with RBC_ERR_ARG; split the queries.  The call is
 * asynchronous: it returns once its work is queued on `stream`.  All calls on one GPU
 * share that GPU's array, so calls that overlap in time there (on different streams or
 * threads) race; calls queued on one stream are ordered.  No scratch word needs to start at zero: each is
 * written before it is read within the call. */
int rbc_brute_k_noalloc(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                        uint32_t ldq, uint32_t d, int metric, uint32_t k, uint32_t *nn_ids,
                        float *nn_dists, uint32_t ldo, void *stream);

/* bruteSearch (brute.cu:43-62 -> nnWrap -> nnKernel kernels.cu:165-222): 1-NN ids and
 * distances; ties go to the lowest index.  nn_dists may be NULL. */
int rbc_brute_search(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, uint32_t *nn_ids, float *nn_dists,
                     int mem, void *stream);

/* bruteRangeCount (brute.cu:19-40 -> rangeCountKernel kernels.cu:426-471): for each
 * query, the number of points with distance < ranges[i]. */
int rbc_brute_range_count(const float *x, uint32_t n, uint32_t ldx, const float *q,
                          uint32_t m, uint32_t ldq, uint32_t d, int metric,
                          const float *ranges, uint32_t *counts, int mem, void *stream);

/* bruteCPU (brute.cu:91-112): the reference's host 1-NN, kept as the host routine it
 * is (single thread, product rounded before the add, strict '<', host pointers). */
int rbc_brute_cpu(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                  uint32_t ldq, uint32_t d, int metric, uint32_t *nn_ids);

/* ------------------------------------------------------- one-shot RBC (reference) -- */

/* buildRBC (rbc.cu:126-208): picks num_reps representatives (rep_ids, or a seeded
 * sample of distinct rows when rep_ids is NULL -- the reference seeds from the clock,
 * utils.cu:17-37), computes each rep's bisection range so that between 0.9 s and s
 * points fall inside (kernels.cu:322-414), and stores the ascending lists of points
 * with distance < range.  Fills *out exactly like the reference: dx, dxMap (num_reps x
 * PAD(s)) and dr on the device, groupCount[PAD(num_reps)] on the host.  x is host
 * memory when mem = RBC_MEM_HOST.  Returns RBC_ERR_NOCONV if a bisection does not
 * settle within 1000 steps (the reference loops forever, todo.txt:3-5). */
int rbc_build(const float *x, uint32_t n, uint32_t d, uint32_t ldx, uint32_t num_reps,
              uint32_t s, const uint32_t *rep_ids, uint64_t seed, int mem, rbcStruct *out);

/* queryRBC (rbc.cu:19-68): one-shot 1-NN (nearest rep, then that rep's list). */
int rbc_query(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs,
              uint32_t *nn_ids, float *nn_dists, int mem, void *stream);

/* kqueryRBC (rbc.cu:72-123): one-shot k-NN, KMAX columns (ld ldo >= KMAX), ascending by
 * (distance, index); unfilled slots hold (MAX_REAL, DUMMY_IDX). */
int rbc_kquery(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs,
               uint32_t *nn_ids, float *nn_dists, uint32_t ldo, int mem, void *stream);

/* The one-shot queries under a general computation plan (compPlan, defs.h:86-92, which
 * initCompPlan builds from a charMatrix, rbc.cu:375-431; queryRBC/kqueryRBC fix it to
 * idIntersection, rbc.cu:46-53 and 99-109).  plan is a host R x R byte matrix with row
 * stride ldp >= R: the queries whose nearest rep is i scan the list of every rep j with
 * plan[i * ldp + j] != 0 (planNNKernel / planKNNKernel, kernels.cu:15-161).  The result
 * is the best (k = 1) or the KMAX best of the union of those lists as a multiset: a point
 * in two planned lists appears twice in a k-NN row, as in the reference.  A group with an
 * empty plan row gets (MAX_REAL, DUMMY_IDX).  fullIntersection (rbc.cu:305-312) is a
 * plan of all ones. */
int rbc_query_plan(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs,
                   const uint8_t *plan, uint32_t ldp, uint32_t *nn_ids, float *nn_dists,
                   int mem, void *stream);
int rbc_kquery_plan(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs,
                    const uint8_t *plan, uint32_t ldp, uint32_t *nn_ids, float *nn_dists,
                    uint32_t ldo, int mem, void *stream);

/* destroyRBC (rbc.cu:363-368). */
int rbc_destroy(rbcStruct *rbcs);

/* The reference's build and query internals (rbc.h:15-26), for callers that assemble a
 * one-shot search step by step.  Device matrices are device pointers, as in the
 * reference; their leading dimensions must be multiples of 4 (the reference's PAD(c) =
 * multiple of 16 always is).  The distance is the process-wide metric (rbc_set_metric).
 *
 * distSubMat (rbc.cu:356-360, over dist1Wrap, kernelWrap.cu:13-34): D[i][j] = distance of
 * rep row start+i of r to point row j of x for i < length, j < nx (FP32, FMA accumulation
 * in coordinate order), D with leading dimension ldd >= nx. */
int rbc_dist_sub(const float *r, uint32_t ldr, uint32_t start, uint32_t length, const float *x,
                 uint32_t nx, uint32_t ldx, uint32_t d, float *D, uint32_t ldd, void *stream);
/* computeReps (rbc.cu:233-247, over nnWrap): nearest of the R reps for each of the m
 * queries (lowest index among equal distances) and its distance.  q and reps are host or
 * device memory per mem_in, rep_ids and dist_to_reps per mem_out (the reference: device
 * in, host out). */
int rbc_compute_reps(const float *q, uint32_t m, uint32_t ldq, const float *reps, uint32_t R,
                     uint32_t ldr, uint32_t d, uint32_t *rep_ids, float *dist_to_reps, int mem_in,
                     int mem_out, void *stream);
/* initCompPlan (rbc.cu:375-431): the device compPlan of a host computation matrix cm
 * (num_reps x num_reps bytes, row stride ldcm; nonzero = planned), the per-rep query
 * counts group_count_q and list lengths group_count_x.  Allocates the four device arrays;
 * rbc_plan_free (freeCompPlan, rbc.cu:435-440) releases them. */
int rbc_plan_init(compPlan *out, const char *cm, uint32_t ldcm, const uint32_t *group_count_q,
                  const uint32_t *group_count_x, uint32_t num_reps);
int rbc_plan_free(compPlan *plan);
/* computeNNs / computeKNNs (rbc.cu:317-351, over planNNKernel / planKNNKernel,
 * kernels.cu:15-161): each query in the device query map qmap (comp_length slots, blocks
 * of 16 slots taking the group of their first slot, DUMMY_IDX = empty slot) scans the
 * lists its group plans: entries [0, groupCountX) of rows qGroupToXGroup of the device
 * xMap (xmap_rows x ldxmap, point ids < nx) over the device database x.  The K best
 * (K = 1: computeNNs; KMAX: computeKNNs) of the union of those lists as a multiset, per
 * query row 0..m-1 (ld ldo; a query in no slot gets (MAX_REAL, DUMMY_IDX)), go to ids /
 * dists in host or device memory per mem_out.  The plan and maps are checked (ranges,
 * each query at most once) before any point is read. */
int rbc_plan_query(const float *x, uint32_t nx, uint32_t ldx, const uint32_t *xmap,
                   uint32_t xmap_rows, uint32_t ldxmap, const float *q, uint32_t m, uint32_t ldq,
                   uint32_t d, const uint32_t *qmap, uint32_t comp_length, const compPlan *plan,
                   uint32_t K, uint32_t *ids, float *dists, uint32_t ldo, int mem_out, void *stream);
/* setupReps (rbc.cu:212-229) with the reps given: copies rows rep_ids[0..R) of the host
 * matrix x into a new zero-padded device matrix (PAD(R) x PAD(d), ld PAD(d)). */
int rbc_setup_reps(const float *x, uint32_t n, uint32_t d, uint32_t ldx, const uint32_t *rep_ids,
                   uint32_t R, matrix *out);

/* ---------------------------------------------------------------- exact RBC ----- */

/* Opaque handle of an exact-search index: representatives, the database permuted
 * into contiguous ownership lists (FP32 and BF16 copies), per-list radii and FP64
 * norms.  No reference counterpart: the reference's exact-search pieces are unused
 * or uncompiled (rbc.cu:250-256 computeRadii, deprecated.cu:354-464 pruneKernel,
 * deprecated.cu:395-428 groupPoints). */
typedef struct rbc_exact rbc_exact;

#define RBC_PREC_FP32 0 /* difference-form FP32 distances on the vector ALUs */
#define RBC_PREC_BF16 1 /* BF16 MFMA distance filter + exact FP32 re-rank (L2, d <= 128,
                           lists of at most 2^25 points) */

/* Build: rep_ids (num_reps distinct rows) or a seeded sample; every point is owned by
 * its nearest rep (lowest rep index on ties); lists are stored contiguously in
 * ascending original-index order. */
int rbc_exact_build(const float *x, uint32_t n, uint32_t d, uint32_t ldx, uint32_t num_reps,
                    int metric, const uint32_t *rep_ids, uint64_t seed, int mem,
                    rbc_exact **out);

typedef struct {
  double pairs_probe;     /* (query, point) distances evaluated by the first-probe pass */
  double pairs_scan;      /* ... by the pruned list scan */
  double pairs_rerank;    /* candidates re-ranked exactly */
  double items_scan;      /* (list, query tile) work items of the scan */
  double lists_scanned;   /* (query, list) pairs that survived pruning */
  double bytes_scan;      /* algorithmic bytes of the scan kernel (list tiles + queries) */
  double guard_fp64;      /* candidates whose order was settled in FP64 */
  double fallbacks;       /* queries re-run because a candidate buffer overflowed */
  double ms_qreps, ms_probe, ms_plan, ms_scan, ms_select, ms_total; /* hipEvent times */
  double flops_scan;      /* MFMA/FMA FLOPs the scan spends per (query, point) pair */
  double sketch;          /* 1 if the scan ran on the projected sketches, else 0 */
} rbc_query_stats;

/* Exact k-NN: ids (original row numbers) and distances of the k nearest points, ascending
 * by (distance, index), ties to the lowest index.  Distances are FP32 re-ranked values
 * within 1e-6 relative of FP64.  precision selects the distance filter.  stats may be
 * NULL. */
int rbc_exact_kquery(rbc_exact *h, const float *q, uint32_t m, uint32_t ldq, uint32_t k,
                     int precision, uint32_t *nn_ids, float *nn_dists, uint32_t ldo, int mem,
                     void *stream, rbc_query_stats *stats);

/* The query-to-representative distances of the exact query's first step (the role of
 * computeReps' distance pass, rbc.cu:233-247): D[i * ldd + j] = distance value (squared for
 * L2) of query i to rep j, as the query computes them for `precision` -- FP32 on the vector
 * ALUs, or split-BF16 MFMA for RBC_PREC_BF16 (L2) -- with |D - exact| <= *rel_err * exact
 * + abs_err[i] (rel_err 0 for BF16; abs_err 0 for FP32; either may be NULL). */
int rbc_exact_rep_dists(rbc_exact *h, const float *q, uint32_t m, uint32_t ldq, int precision,
                        float *D, uint32_t ldd, float *abs_err, double *rel_err, int mem,
                        void *stream);

typedef struct {
  uint32_t n, d, ld, num_reps, metric;
  uint32_t max_list, min_list;
  size_t device_bytes;
} rbc_exact_info;

int rbc_exact_get_info(const rbc_exact *h, rbc_exact_info *info);

/* Copies the build's structure to host arrays (any may be NULL): rep_ids[num_reps],
 * list_offsets[num_reps+1], perm[n] (original id of permuted row), owner[n] (by original
 * id), radii[num_reps] (max FP64 distance of a list's points to its rep, rounded up). */
int rbc_exact_get_lists(const rbc_exact *h, uint32_t *rep_ids, uint32_t *list_offsets,
                        uint32_t *perm, uint32_t *owner, float *radii);

int rbc_exact_destroy(rbc_exact *h);

/* Projected lower-bound filter of the BF16 k-query (L2, 32 < d <= 128).  The build keeps
 * P, the 28 leading principal directions of a 65,536-row sample of the database
 * (orthonormal rows, FP32), and for every point the BF16 coordinates of P (x - mu) with an
 * error bound.  Since |P (q - x)| <= |q - x|, the list scan can test the 28-coordinate
 * projections (one 16x16x32 MFMA per 16 x 16 pairs, norm and threshold carried in four
 * extra coordinates) instead of the full rows; the exact FP32 re-rank decides the results,
 * which are the same either way.  RBC_SKETCH_AUTO (the default) uses it when twice the
 * data's variance outside P is at most 5% of the mean squared point-to-representative
 * distance (the bound then stays tight); ON and OFF force it.  No reference counterpart. */
#define RBC_SKETCH_AUTO 0
#define RBC_SKETCH_OFF 1
#define RBC_SKETCH_ON 2
int rbc_exact_set_sketch(rbc_exact *h, int mode);

typedef struct {
  uint32_t dims;          /* projection rows (0: the index has no sketch) */
  uint32_t active;        /* 1 if BF16 k-queries scan the sketches */
  double norm_bound;      /* s_P >= |P|_2 */
  double resid_var;       /* sample variance outside the projection (trace - kept) */
  double total_var;       /* sample variance (trace of the covariance) */
  double mean_owner_d2;   /* sample mean of |x - owner rep|^2 */
} rbc_sketch_info;

int rbc_exact_sketch_info(const rbc_exact *h, rbc_sketch_info *info);

/* Sketches of m rows (host or device, leading dimension ldx): out[i * 32 + j] = BF16 bits
 * of y~_j = bf16((P (x_i - mu))_j) for j < dims (the remaining coordinates are the query
 * form's 1, 1, 0, 0), err[i] >= |y~ - P (x_i - mu)| (P and mu the stored FP32 values), n2[i]
 * = |y~|^2; any output may be NULL.  The rows of the index itself give its stored sketch
 * coordinates.  For tests of the bound. */
int rbc_exact_sketch_rows(rbc_exact *h, const float *x, uint32_t m, uint32_t ldx, int mem,
                          uint16_t *out, float *err, float *n2);
/* P (dims x d, row-major) and mu (d) as stored (FP32, host buffers; either may be NULL). */
int rbc_exact_sketch_basis(const rbc_exact *h, float *P, float *mu);

/* Build once, query many (the reference's readme.txt:62-65; the reference has no file
 * format for an index): writes every array of the index to path, and loads such a file
 * into a new index that answers exactly like the saved one.  No counterpart in the
 * reference. */
int rbc_exact_save(const rbc_exact *h, const char *path);
int rbc_exact_load(const char *path, rbc_exact **out);

/* ------------------------------------------------------- learned list proposal -- */

/* A small MLP d_in -> h1 -> h2 -> n_out (GELU, FP16 weights and activations, FP32
 * accumulation) trained on the GPU by the library's own kernels.  For an index (n_out =
 * num_reps) it scores the ownership lists by how likely they hold a query's nearest
 * neighbours; trained on the lists that hold the exact k NNs of training queries (soft
 * targets), its ordering decides which lists the exact query probes first.  No reference
 * counterpart. */
typedef struct rbc_mlp rbc_mlp;

int rbc_mlp_create(uint32_t d_in, uint32_t h1, uint32_t h2, uint32_t n_out, uint64_t seed,
                   rbc_mlp **out);
int rbc_mlp_destroy(rbc_mlp *m);
int rbc_mlp_num_params(const rbc_mlp *m, size_t *count);
/* FP32 master parameters, concatenated: W1[h1][d_in], b1[h1], W2[h2][h1], b2[h2],
 * W3[n_out][h2], b3[n_out] (host buffers). */
int rbc_mlp_get_params(const rbc_mlp *m, float *params);
int rbc_mlp_set_params(rbc_mlp *m, const float *params);
/* Proposal network to / from a file: layer sizes, parameters and Adam state. */
int rbc_mlp_save(const rbc_mlp *m, const char *path);
int rbc_mlp_load(const char *path, rbc_mlp **out);
/* Mean softmax cross-entropy of one batch (host x, labels) and its gradient w.r.t. the
 * parameters (host, same layout as the parameters); no update. */
int rbc_mlp_loss_grad(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx,
                      const uint32_t *labels, float *loss, float *grads);
/* The same with labels_per_row (1..64) labels per row, labels[rows][labels_per_row]: row
 * i's target is uniform over its labels below n_out (a label >= n_out is no label; a row
 * without any contributes nothing).  labels_per_row = 1 is rbc_mlp_loss_grad. */
int rbc_mlp_loss_grad_multi(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx,
                            const uint32_t *labels, uint32_t labels_per_row, float *loss,
                            float *grads);
/* Logits (host, rows x n_out) of host inputs. */
int rbc_mlp_logits(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, float *logits);
/* The proposal of each host input row: the index of its largest logit, lowest index
 * among equal logits (UINT_MAX if every logit is NaN) -- the list the exact query adds
 * to its probe. */
int rbc_mlp_propose(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, uint32_t *out);
/* Adam training (beta 0.9/0.999) for `epochs` passes over shuffled batches; epoch_loss
 * (may be NULL) receives the loss of each epoch's last batch. */
int rbc_mlp_train(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, const uint32_t *labels,
                  int mem, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                  float *epoch_loss);
/* Training on soft targets (labels[rows][labels_per_row], as rbc_mlp_loss_grad_multi),
 * otherwise as rbc_mlp_train; labels_per_row = 1 is rbc_mlp_train. */
int rbc_mlp_train_multi(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx,
                        const uint32_t *labels, uint32_t labels_per_row, int mem, uint32_t epochs,
                        uint32_t batch, float lr, uint64_t seed, float *epoch_loss);
/* The P (1..16) largest logits of each host input row, descending (lowest index among
 * equal logits), as lists[rows][P], and in counts[rows] how many of them it takes for
 * their softmax probabilities to sum to at least `mass` (0 if all P fall short) -- the
 * probe set the exact query takes from the network. */
int rbc_mlp_propose_lists(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, uint32_t P,
                          float mass, uint32_t *lists, uint32_t *counts);
/* Use (or, with NULL, stop using) the network to choose the lists the exact k-query probes
 * first.  BF16 filter: a query probes the network's best lists, as many as it takes for
 * their softmax mass to reach the proposal mass (default 0.9), at most 8; a query whose
 * 8 best fall short probes the lists of its 8 nearest representatives, as without the
 * network.  FP32 filter: the network's top-1 list is a second probe.  Results do not
 * depend on it (the probe only sets the pruning bound).  The network is not owned by the
 * index. */
int rbc_exact_set_proposal(rbc_exact *h, rbc_mlp *mlp);
/* Search parameters of the BF16 k-query (no effect on results): the number of lists each
 * query probes for its first bound (1..8, default 8) and the number of nearby reps whose
 * Voronoi planes bound the lists in the plan besides the nearest rep (0..12; default 8
 * when the scan runs on sketches, 12 otherwise). */
int rbc_exact_set_search_params(rbc_exact *h, uint32_t probe_lists, uint32_t plane_reps);
/* The proposal mass of the BF16 probe set, in (0, 1]. */
int rbc_exact_set_proposal_mass(rbc_exact *h, float mass);

/* ------------------------------------------------------------ batched exact search -- */

/* Many independent exact indexes (each its own database, d <= 128, num_reps <= 1024),
 * queried together: one persistent kernel walks every index's query tiles, captured in
 * a hipGraph that is replayed while the batch shape and (device) buffers stay the same.
 * The query-tile size is autotuned on the first call for a batch shape.  No reference
 * counterpart. */
typedef struct rbc_batched rbc_batched;

typedef struct {
  double ms_graph;          /* hipEvent time of the graph launch */
  uint32_t tile_queries;    /* queries per tile (16 per wave for L2; else autotuned 16/32/64) */
  uint32_t graph_replayed;  /* 1 if the call replayed a captured graph */
  uint32_t uncertified;     /* queries re-run through the single-index path */
  uint32_t blocks_per_cu;   /* L2: autotuned resident workgroups per CU of the wave kernel */
  double tune_ms[3];        /* autotuner timings (first call): L2 a quarter, half and all of the
                               workgroups per CU that fit; otherwise 16-, 32-, 64-query tiles */
  double guard_fp64;        /* candidates ordered in FP64 by the near-tie guard */
  double pairs_scan;        /* (query, point) distances the persistent kernel evaluated */
  double flops;             /* 2 d per pair scanned plus 2 d per (query, rep) distance */
  double steps;             /* 64-point list steps the persistent kernel ran (all tiles) */
  double longest_tile_ms;   /* longest single tile (device wall clock), the launch's floor */
} rbc_batched_stats;

/* count indexes; x[i] has n[i] rows of d[i] floats with leading dimension ldx[i]; index
 * i uses num_reps[i] representatives drawn from seed + i. */
int rbc_batched_build(uint32_t count, const float *const *x, const uint32_t *n, const uint32_t *d,
                      const uint32_t *ldx, const uint32_t *num_reps, int metric, uint64_t seed,
                      int mem, rbc_batched **out);
/* k-NN (k <= 56) of m[i] queries q[i] (leading dimension ldq[i]) against index i, for all
 * i; results as rbc_exact_kquery into nn_ids[i], nn_dists[i] (leading dimension ldo[i]).
 * Runs on the handle's own stream (`stream` is not used). */
int rbc_batched_kquery(rbc_batched *h, const float *const *q, const uint32_t *m,
                       const uint32_t *ldq, uint32_t k, uint32_t *const *nn_ids,
                       float *const *nn_dists, const uint32_t *ldo, int mem, void *stream,
                       rbc_batched_stats *stats);
/* The exact index of instance i (owned by h: valid until rbc_batched_destroy), for
 * rbc_exact_get_info / rbc_exact_get_lists. */
int rbc_batched_get_index(const rbc_batched *h, uint32_t i, const rbc_exact **out);
int rbc_batched_destroy(rbc_batched *h);

#ifdef __cplusplus
}
#endif
#endif /* RBC_HIP_H */
