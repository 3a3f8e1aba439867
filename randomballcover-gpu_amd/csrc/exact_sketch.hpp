// Projected lower-bound filter of the BF16 exact query (exact_sketch.hip): shapes, the
// scan's arguments and the host entry points.
#pragma once

#include <cfloat>
#include <cmath>

#include "exact.hpp"

namespace rbc {

constexpr int kSkDims = 28;    // projection rows (principal directions)
constexpr int kSkW = 32;       // BF16 coordinates per sketch row (28 + 4 carrying norm and threshold)
constexpr int kSkMaxD = 128;   // largest d with a sketch (the BF16 filter's limit)
constexpr uint32_t kSkScanQ = 1024;  // queries per scan item (8 waves x 8 blocks of 16)
// RBC_SKETCH_AUTO uses the sketch when twice the variance outside the projection is at
// most this fraction of the mean squared point-to-representative distance.
constexpr double kSkAutoFrac = 0.05;

constexpr uint16_t kBf16One = 0x3F80, kBf16NegInf = 0xFF80, kBf16PosInf = 0x7F80;

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
// Smallest BF16 >= f (f not NaN).
__device__ __forceinline__ uint16_t bf16_up(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0xFFFFu) == 0u || f < 0.f) return (uint16_t)(u >> 16);  // exact, or truncation towards zero
  return (uint16_t)((u >> 16) + 1u);                              // may round up to +inf
}
// Largest BF16 <= f for f >= 0.
__device__ __forceinline__ uint16_t bf16_down_pos(float f) { return (uint16_t)(__float_as_uint(f) >> 16); }
__device__ __forceinline__ float f32_down(double v) {
  float f = (float)v;
  if ((double)f > v) f = nextafterf(f, -INFINITY);
  return f;
}
__device__ __forceinline__ float f32_up(double v) {
  if (v >= (double)FLT_MAX) return INFINITY;
  float f = (float)v;
  if ((double)f < v) f = nextafterf(f, INFINITY);
  return f;
}

// Sketch threshold of a (query, list) pair, packed as (s_lo << 16) | s_hi with s_hi +
// s_lo >= s, s = ((G + eta)^2 - nq) / 2 + margin rounded up.  The MFMA adds the 32
// products y~_q,i y~_x,i, -n_hi, -n_lo, s_hi, s_lo exactly and accumulates in FP32: the
// error is at most c_gemm (gamma_32 + 6u') times their absolute sum, <= |y~_q||y~_x| +
// |y~_x|^2 / 2 + |s| (1 + 2^-7) <= nq / 2 + nxm + 1.01 |s|, which the margin c_gemm (2 |s0|
// + nq + nxm) + u' (nq + nxm) covers (|s| exceeds |s0| only by the margin itself).
__device__ __forceinline__ uint32_t sk_threshold_pack(double G, float qe, float qn, float lqe, float lnm, double c_gemm) {
  const double eta = (double)qe + (double)lqe, nq = (double)qn, nxm = (double)lnm;
  const double T0 = (G + eta) * (G + eta);
  const double s0 = 0.5 * (T0 - nq);
  const double up = 0x1p-23;
  const double margin = c_gemm * (2.0 * fabs(s0) + nq + nxm) + up * (nq + nxm) + 1e-30;
  const float hs = f32_up((s0 + margin) * (1.0 + 1e-9));
  if (hs >= INFINITY) return (uint32_t)kBf16PosInf;
  const uint16_t sh = bf16_up(hs);
  const float rem = hs - bf16_val(sh);  // exact (Sterbenz): |sh| within 2^-7 of |hs|
  const uint16_t sl = bf16_up(rem);
  return ((uint32_t)sl << 16) | sh;
}

struct SkScanArgs {
  const ScanItem *items;
  const uint32_t *bounds;  // [kXCD + 1]: items [bounds[x], bounds[x+1]) run on XCD x
  uint32_t *next;          // [kXCD]
  const uint32_t *qids;    // list-major (query, list) pairs: query ids
  const uint32_t *qthr;    // per query: packed emission threshold (sk_threshold_pack)
  const uint16_t *qs;      // query sketch rows [m][kSkW]
  const uint16_t *xs;      // permuted database sketch rows [n][kSkW]
  const uint32_t *perm;
  uint32_t n;
  uint32_t *cand_i;
  uint32_t *cand_n;
  uint32_t cap;
};

struct SkProbeArgs {
  const ScanItem *items;
  uint32_t nitems;
  const uint32_t *qids;    // list-major probe entries (entry e belongs to query e / P)
  const uint16_t *qs;      // query sketch rows
  const float *qn;         // |y~_q|^2
  const uint16_t *xs;      // permuted database sketch rows
  const uint32_t *perm;
  uint32_t n, Kp, P;
  uint32_t idx_bits;       // low key bits holding a point's slot (4 per 16-point block)
  uint32_t *out_i;         // [entries][Kp] original ids, ascending approximate distance
  float *out_v;            // [entries][Kp] approximate squared sketch distances (up to a per-query constant)
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
// The probe on sketches: for every entry (query, list) of items (tile kBf16ProbeQ), its
// Kp (<= 16) closest points by sketch distance, ascending, as bf16_probe returns them (the
// entries of Kp <= 8 come from 2 keys per lane, of more from 4).
int sketch_probe(rbc_exact *h, const uint16_t *qs, const float *qn, const uint32_t *entries, uint32_t P,
                 const ScanItem *items, uint32_t nitems, uint32_t Kp, uint32_t *out_i, float *out_v, hipStream_t st);
// The list scan on sketches: items from make_bucket_items over list_qoff with tile
// kSkScanQ; sched holds 2 kXCD + 1 uint32 of device scratch.  The emission threshold of a
// query is computed once per call (with the index-wide largest point error and norm);
// qthr is device scratch of m uint32 (one per query of the call).
int sketch_scan(rbc_exact *h, uint32_t m, const uint16_t *qs, const float *qe, const float *qn, const double *gam,
                uint32_t *qthr, const uint32_t *qlist, const uint32_t *list_qoff, const ScanItem *items,
                uint32_t nitems, uint32_t *sched, uint32_t *cand_i, uint32_t *cand_n, uint32_t cap, hipStream_t st);

}  // namespace rbc
