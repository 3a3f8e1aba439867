// Device helpers of the exact re-rank shared by the single-index and batched queries:
// tree-ordered FP32 distances, the FP64 distance in the oracle's order, and the near-tie
// guard that orders adjacent results whose FP32 error intervals overlap by FP64.
#pragma once

#include "rbc_device.hpp"

namespace rbc {

// Sorts (run, key64, id) triples held one per lane ascending (run, key, id), carrying
// a float payload.
__device__ void wave_sort_runs(uint32_t &run, double &key, uint32_t &id, float &pay) {
  const int l = lane_id();
#pragma unroll
  for (int k = 2; k <= kWave; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint32_t pr = xor_lane(run, j);
      double pk = xor_lane(key, j);
      uint32_t pi = xor_lane(id, j);
      float pp = xor_lane(pay, j);
      bool p_less = pr < run || (pr == run && (pk < key || (pk == key && pi < id)));
      bool m_less = run < pr || (run == pr && (key < pk || (key == pk && id < pi)));
      bool keep_min = ((l & j) == 0) == ((l & k) == 0);
      if (keep_min ? p_less : m_less) {
        run = pr;
        key = pk;
        id = pi;
        pay = pp;
      }
    }
}

template <int METRIC>
__device__ double fp64_dist(const float *xrow, const float *qrow, uint32_t d) {
  double acc = 0.0;
  for (uint32_t c = 0; c < d; ++c) {
    double t = __dsub_rn((double)xrow[c], (double)qrow[c]);
    acc = __dadd_rn(acc, METRIC == 0 ? __dmul_rn(t, t) : fabs(t));
  }
  return acc;
}

// Lower bound (metric distance) on d(q, x) for every x in list j.
__device__ __forceinline__ double list_lower_bound(double Dqj, double Dq0, double rr_j0, double psi,
                                                   double e, bool same, int metric) {
  double lb;
  if (metric == 0) {
    double dlo = sqrt(fmax(Dqj * (1.0 - e), 0.0));
    lb = dlo - psi;
    if (!same && rr_j0 > 0.0) {
      // Voronoi plane between r_j and the query's nearest rep r_0: every x owned by r_j
      // satisfies |x-r_j|^2 <= |x-r_0|^2 + 2e psi^2 (the owner test ran in FP32), so
      // d(q,x) >= (|q-r_j|^2 - |q-r_0|^2 - 2e psi^2) / (2 |r_j - r_0|).
      double num = Dqj * (1.0 - e) - Dq0 * (1.0 + e) - 2.0 * e * psi * psi;
      double den = 2.0 * sqrt(rr_j0 * (1.0 + e));
      if (num > 0.0) lb = fmax(lb, num / den);
    }
  } else {
    lb = Dqj * (1.0 - e) - psi;
  }
  return lb * (1.0 - 1e-12);
}

// Re-ranks candidates 0..total-1 of one query (cand(e) -> original id, kDummy to skip)
// with tree-ordered FP32 distances and keeps the 64 best in (tv, ti), ascending by
// (distance, id), one per lane.  Eight candidate rows are loaded at a time.
template <int METRIC, int NPL, class CandFn>
__device__ void rerank_candidates(float &tv, uint32_t &ti, uint32_t total, CandFn cand, const float *xp,
                                  size_t ldx, const uint32_t *inv, const float (&qv)[NPL], uint32_t d) {
  const int l = lane_id();
  for (uint32_t b0 = 0; b0 < total; b0 += kWave) {
    const uint32_t e = b0 + l;
    uint32_t myid = e < total ? cand(e) : kDummy;
    const uint32_t myrow = myid != kDummy ? inv[myid] : 0u;
    float myv = FLT_MAX;
    const uint32_t lim = min((uint32_t)kWave, total - b0);
    for (uint32_t cc = 0; cc < lim; cc += 8) {
      float t[8][NPL];
      bool ok[8];
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        const uint32_t r = __shfl(myrow, cc + g, kWave);
        ok[g] = cc + g < lim && __shfl(myid, cc + g, kWave) != kDummy;
        const float *xrow = xp + size_t(r) * ldx;
#pragma unroll
        for (int j = 0; j < NPL; ++j) {
          const uint32_t c = l + kWave * j;
          t[g][j] = ok[g] && c < d ? tree_term<METRIC>(xrow[c], qv[j]) : 0.f;
        }
      }
#pragma unroll
      for (int g = 0; g < 8; ++g) {
#pragma unroll
        for (int s = NPL / 2; s >= 1; s >>= 1)
#pragma unroll
          for (int j = 0; j < s; ++j) t[g][j] = __fadd_rn(t[g][j], t[g][j + s]);
        const float v = wave_tree_sum(t[g][0]);
        if (ok[g] && l == (int)(cc + g)) myv = v;
      }
    }
    wave_bitonic_sort(myv, myid);
    wave_merge_lower(tv, ti, myv, myid);
  }
}

// (tv, ti): the 64 best candidates of a query, ascending by (FP32 tree value, id), one
// per lane; `total` = number of candidates seen.  Adjacent entries whose error intervals
// [v(1-e), v(1+e)] overlap form runs; runs that start inside the first k slots are
// re-sorted by their FP64 distances.  The reported values stay the FP32 ones.
template <int METRIC>
__device__ void near_tie_guard(float &tv, uint32_t &ti, uint32_t total, const float *xp, size_t ldx,
                               const uint32_t *inv, const float *qrow, uint32_t d, uint32_t k,
                               double e_tree, unsigned long long *stats) {
  const int l = lane_id();
  // Near-tie guard: adjacent results whose error intervals overlap are ordered by their
  // FP64 distances (computed exactly as the FP64 oracle does).
  const float er = (float)e_tree;
  float nv = __shfl_down(tv, 1, kWave);
  bool ov = l < kWave - 1 && tv < FLT_MAX && nv < FLT_MAX && nv * (1.f - er) <= tv * (1.f + er);
  bool ov_prev = __shfl_up(ov, 1, kWave) && l > 0;
  uint64_t bmask = __ballot(!ov_prev);  // run starts
  uint64_t upto = bmask & (l == 63 ? ~0ull : ((2ull << l) - 1));
  int start = 63 - __clzll(upto);
  uint32_t run = (uint32_t)__popcll(upto);
  bool amb = (ov || ov_prev) && start < (int)k;
  double key = (double)tv;
  if (__any(amb)) {
    if (amb) {
      const float *xrow = xp + size_t(inv[ti]) * ldx;
      key = fp64_dist<METRIC>(xrow, qrow, d);
    }
    uint32_t cnt = (uint32_t)__popcll(__ballot(amb));
    if (l == 0) atomicAdd(&stats[0], (unsigned long long)cnt);
    // A near-tie run that reaches the last kept slot could continue past it.
    if (__shfl(amb ? 1 : 0, kWave - 1, kWave) && total > (uint32_t)kWave && l == 0)
      atomicAdd(&stats[2], 1ull);
    // Reported distances stay the FP32 re-ranked values, carried with their ids.
    wave_sort_runs(run, key, ti, tv);
  }
}

}  // namespace rbc
