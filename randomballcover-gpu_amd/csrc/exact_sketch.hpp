// Projected lower-bound filter of the BF16 exact query (exact_sketch.hip): shapes, the
// scan's arguments and the host entry points.
#pragma once

#include "exact.hpp"

namespace rbc {

constexpr int kSkDims = 28;    // projection rows (principal directions)
constexpr int kSkW = 32;       // BF16 coordinates per sketch row (28 + 4 carrying norm and threshold)
constexpr int kSkMaxD = 128;   // largest d with a sketch (the BF16 filter's limit)
constexpr uint32_t kSkScanQ = 1024;  // queries per scan item (4 waves x 16 blocks of 16)
// RBC_SKETCH_AUTO uses the sketch when twice the variance outside the projection is at
// most this fraction of the mean squared point-to-representative distance.
constexpr double kSkAutoFrac = 0.05;

struct SkScanArgs {
  const ScanItem *items;
  const uint32_t *bounds;  // [kXCD + 1]: items [bounds[x], bounds[x+1]) run on XCD x
  uint32_t *next;          // [kXCD]
  const uint32_t *qids;    // list-major (query, list) pairs: query ids
  const uint16_t *qs;      // query sketch rows [m][kSkW]
  const float *qn;         // |y~_q|^2
  const float *qe;         // e_q
  const double *gam;       // per query: upper bound on the k-th NN distance
  double sP;               // bound on |P|_2
  double c_gemm;           // gamma_32 + 6u'
  const float *lqe;        // per list: max e_x
  const float *lnm;        // per list: max |y~_x|^2 (rounded up)
  const uint16_t *xs;      // permuted database sketch rows [n][kSkW]
  const uint32_t *perm;
  uint32_t n;
  uint32_t *cand_i;
  uint32_t *cand_n;
  uint32_t cap;
};

// Builds the projection, the database sketch rows and their per-list maxima from the
// index's permuted rows (h->xp, lists, reps); no sketch (sk_dims = 0) unless L2 and
// 32 < d <= kSkMaxD.
int sketch_build(rbc_exact *h, hipStream_t st);
// Sets h->sk_mode and h->sk_active (RBC_SKETCH_*).
void sketch_set_mode(rbc_exact *h, int mode);
// gamma_{d+1} (u = 2^-24) with 1% headroom: the projection's per-coordinate error factor.
double sketch_proj_err(uint32_t d);
// gamma_32 + 6u' (u' = 2^-23) of the scan's FP32 accumulation.
double sketch_gemm_err();
// Query sketch rows qs[m][kSkW] (coordinates 30, 31 left 0 for the scan), error bounds
// qe and |y~_q|^2 (qn).
int sketch_queries(const rbc_exact *h, const float *q, size_t ldq, uint32_t m, uint16_t *qs, float *qe, float *qn,
                   hipStream_t st);
// The list scan on sketches: items from make_bucket_items over list_qoff with tile
// kSkScanQ; sched holds 2 kXCD + 1 uint32 of device scratch.
int sketch_scan(rbc_exact *h, const uint16_t *qs, const float *qn, const float *qe, const double *gam,
                const uint32_t *qlist, const uint32_t *list_qoff, const ScanItem *items, uint32_t nitems,
                uint32_t *sched, uint32_t *cand_i, uint32_t *cand_n, uint32_t cap, hipStream_t st);

}  // namespace rbc
