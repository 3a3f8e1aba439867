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
  const uint32_t *probe_i;  // [m][Kp] probe candidates (original ids)
  const uint32_t *cert;     // per query: probe candidates complete (else ignored)
  const uint32_t *cand_i;   // [m][cap] scan candidates
  const uint32_t *cand_n;
  uint32_t cap;
  double e_tree;  // relative error bound of the tree-ordered FP32 re-rank
  float *out_d;
  uint32_t *out_i;
  uint32_t ldo;
  unsigned long long *stats;  // [0] FP64-guarded slots, [1] re-ranked, [2] runs at the edge
};

int launch_select(const SelectArgs &a, int metric, hipStream_t st);

// BF16 MFMA stages (exact_bf16.hip).
int bf16_qreps(rbc_exact *h, const float *q, size_t ldq, uint32_t m, float *Dqr, hipStream_t st);
int bf16_probe(rbc_exact *h, const float *q, size_t ldq, const uint32_t *qorder, const ScanItem *items,
               uint32_t nitems, uint32_t Kp, float *probe_d, uint32_t *probe_i, hipStream_t st);
int bf16_bounds(rbc_exact *h, const float *q, size_t ldq, uint32_t m, uint32_t k, uint32_t Kp,
                const float *probe_d, const float *rep_kth, const uint32_t *near, float *tau,
                double *gam, uint32_t *cert, hipStream_t st);
int bf16_scan(rbc_exact *h, const float *q, size_t ldq, const uint32_t *qlist, const ScanItem *items,
              uint32_t nitems, const float *tau, float *cand_d, uint32_t *cand_i, uint32_t *cand_n,
              uint32_t cap, hipStream_t st);
double bf16_plan_err(const rbc_exact *h);

}  // namespace rbc
