// Shared declarations of the exact-query pipeline (FP32 and BF16 stages).
#pragma once

#include "exact.hpp"

namespace rbc {

struct SelectArgs {
  const float *q;
  size_t ldq;
  const float *xp;
  size_t ldx;
  const uint32_t *inv;
  uint32_t m, d, k, Kp;
  uint32_t n;                 // database rows (bounds checks)
  const uint32_t *probe_i;   // [m][Kp] first-probe candidates (original ids)
  // With probe_d: probe_i/probe_d are the probe's K' distinct points re-ranked exactly as
  // select re-ranks (tree-ordered FP32), ascending; select starts from them and re-ranks
  // only the scan candidates not among them.  NULL: probe_i are candidates like the scan's.
  const float *probe_d;
  const uint32_t *probe_i1;  // [m][Kp] second-probe candidates, or NULL
  const uint32_t *key1;      // second-probe list per query (R = none)
  uint32_t R;
  const uint32_t *cert;      // per query: bit s = probe slot s complete (else ignored)
  const uint32_t *cand_i;   // [m][cap] scan candidates
  const uint32_t *cand_n;
  uint32_t cap;
  double e_tree;  // relative error bound of the tree-ordered FP32 re-rank
  float *out_d;
  uint32_t *out_i;
  uint32_t ldo;
  unsigned long long *stats;  // [0] FP64-guarded slots, [2] near-tie runs reaching the last kept slot
};

int launch_select(const SelectArgs &a, int metric, hipStream_t st);

// BF16 MFMA stages (exact_bf16.hip).
// With qn2 (and qerr), also ||q||^2 and the absolute error bound of the query's
// split-BF16 rep distances (bf16_qr_dist); with qlo, the low halves of the split operands
// (m x ldb, beside qb's high halves).
int bf16_prepare_queries(rbc_exact *h, const float *q, size_t ldq, uint32_t m, uint16_t *qb, float *qn,
                         float *qe, hipStream_t st, float *qn2 = nullptr, float *qerr = nullptr,
                         uint16_t *qlo = nullptr);
// Relative constant c of the split-BF16 dot products: |q.r~ - q.r| <= c |q| |r|.
double qr_split_err(const rbc_exact *h);
// L2 query-rep distances D~ (m x R) on MFMA from split BF16 operands (qhi = qb, qlo and
// qn2 from bf16_prepare_queries); |D~ - D| <= qerr[q] for every rep.
int bf16_qr_dist(const rbc_exact *h, const uint16_t *qhi, const uint16_t *qlo, uint32_t m, const float *qn2, float *D,
                 hipStream_t st);
// Probe entries e (query e / P): for each, the K' smallest approximate distances of its
// item's list, ascending, as ids out_i[e*Kp..] and values out_v[e*Kp..].
int bf16_probe(rbc_exact *h, const uint16_t *qb, const float *qn, const uint32_t *entries, uint32_t P,
               const ScanItem *items, uint32_t nitems, uint32_t Kp, uint32_t *out_i, float *out_v,
               hipStream_t st);
// Lists probed per query by the BF16 probe.
constexpr uint32_t kProbeLists = 8;
// Reps per query whose Voronoi planes bound the lists in the exact plan (<= 64).
constexpr uint32_t kPlaneReps = 12;
// Queries per work item of the BF16 scan (its items must be cut at this size).
constexpr uint32_t kBf16ScanQ = 512;
// Queries per work item of the BF16 probe.
constexpr uint32_t kBf16ProbeQ = 64;
// Candidates the BF16 probe finds per (query, list) entry (4 lanes x 4 kept keys).
constexpr uint32_t kBf16ProbeKeep = 16;
// Relative error constant of the BF16 GEMM distance: gamma_K + 6u', u' = 2^-23.
double bf16_gemm_err(const rbc_exact *h);
// Scan items must come from make_bucket_items over list_qoff (R + 1 offsets of the
// list-major query ids) with tile kBf16ScanQ; sched holds 17 uint32 of device scratch.
int bf16_scan(rbc_exact *h, const uint16_t *qb, const float *qn, const float *qe, const double *gam,
              const uint32_t *qlist, const uint32_t *list_qoff, const ScanItem *items, uint32_t nitems,
              uint32_t *sched, float *cand_d, uint32_t *cand_i, uint32_t *cand_n, uint32_t cap,
              hipStream_t st);

// Emission threshold of the BF16 list scan for one (query, list) pair.  The scan's
// accumulator starts at -|x~|^2/2 (the stored FP32 norm halved, exactly) and adds the K
// products q~_i x~_i; it emits x when the accumulator reaches -s, with the returned s =
// (T - |q~|^2)/2 + margin rounded up, T = (G + eta)^2, eta = |q - q~| + max_list |x - x~|.
// Any x with |q - x| <= G has |q~ - x~| <= G + eta, i.e. q~.x~ - |x~|^2/2 >= -(T - |q~|^2)/2
// in exact arithmetic.  The accumulator's error is at most c_gemm (gamma_K + 6u' of the
// absolute terms) times |x~|^2/2 + |q~||x~| <= nxm + nq/2, and the stored norm is within
// u' nxm of |x~|^2 (u' = 2^-23); the margin c_gemm (|s0| + nq + nxm) + u' (nq + nxm) covers
// both with room to spare.
__device__ __forceinline__ float bf16_threshold(double G, float qe, float qn, float lqe, float lnm,
                                                double c_gemm) {
  const double eta = (double)qe + (double)lqe;
  const double nq = (double)qn;
  const double nxm = (double)lnm;
  const double T0 = (G + eta) * (G + eta);
  const double s0 = 0.5 * (T0 - nq);
  const double up = 0x1p-23;
  const double margin = c_gemm * (fabs(s0) + nq + nxm) + up * (nq + nxm) + 1e-30;
  const double hh = (s0 + margin) * (1.0 + 1e-9);
  float hs = hh >= (double)FLT_MAX ? INFINITY : (float)hh;
  if ((double)hs < hh) hs = nextafterf(hs, INFINITY);
  return hs;
}

}  // namespace rbc
