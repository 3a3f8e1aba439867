// Exact k-NN query over an rbc_exact index.
//
// Pipeline (one stream; the proposal network's inference runs beside step 1 on a second):
//  1. q->reps   : distances of every query to every representative: FP32 on the vector
//                 ALUs (relative error e), or, for the BF16 filter, split-BF16 MFMA tiles
//                 with a per-query absolute error bound (exact_bf16.hip, qr_mfma).
//  2. row stats : nearest rep of each query and an upper bound on its k-th smallest rep
//                 distance (reps are database points, so that bounds the k-th NN distance);
//                 BF16: also the reps of its smallest lane minima (probe lists, plane reps).
//  3. probe     : an upper bound gamma on the k-th NN distance (and a cut-off tau for
//                 FP32 candidates) from K' = k+8 database points near the query.  FP32:
//                 queries grouped by nearest rep scan that rep's list (and the proposal's
//                 list) keeping the best K'; a list whose probe is complete is not scanned
//                 again.  BF16: each query's 8 lists of smallest lane-minimum rep distance
//                 (one slot going to the proposal's list) are scanned in one pass of MFMA
//                 tiles; the query's best K' approximate pairs are re-ranked in FP32.
//  4. plan      : per (query, list) lower bounds on the distance to any point of the
//                 list -- the triangle inequality with the list radius
//                 (d(q,r_j) - psi_j) and, for L2, the Voronoi planes between the list's
//                 rep and the query's nearest rep (BF16: and its plane reps) -- with
//                 explicit FP32 error bounds and directed rounding; lists whose bound
//                 exceeds gamma are pruned.  The surviving (list, query) pairs are
//                 compacted list-major.
//  5. scan      : every (list, query tile) item computes all distances and appends the
//                 ones that can be within gamma to per-query candidate buffers.
//  6. select    : one wave per query re-ranks the candidates with tree-ordered FP32 sums,
//                 sorts by (distance, id) with wave-level bitonic merges, settles
//                 near-ties in FP64 and writes the k results.
// The BF16 variant (precision RBC_PREC_BF16) runs steps 3 and 5 on BF16 MFMA tiles whose
// results carry explicit error bounds (exact_bf16.hip).
#include <algorithm>
#include <chrono>
#include <cmath>

#include "exact.hpp"
#include "exact_device.hpp"
#include "exact_query.hpp"
#include "exact_sketch.hpp"

namespace rbc {

namespace {

// ---- 1. q -> reps ---------------------------------------------------------------------
// All query-rep distances (metric value, FP32): per pair, acc = 0 and one fma_term per
// coordinate in coordinate order -- the values of the tile engine's EpiDense, so the
// error bound fma_rel_err(dpad) holds.  128 x 128 tiles, 8 x 8 per thread (rows
// ty*4.. and 64+ty*4.. of the query tile, columns tx*4.. and 64+tx*4.. of the rep tile),
// 32 coordinates at a time staged transposed in LDS.  Staging: 8 threads read one row's
// 32 coordinates as 16-byte loads (a 128-byte run; the query tile falls back to 4-byte
// loads when its rows are not 16-byte aligned), and write them transposed into rows
// padded to 132 floats (at most 2-way bank conflicts on the stores).
constexpr int kQRT = 128, kQRC = 32, kQRP = kQRT + 4;

// Coordinates c..c+3 of a row (zero from d on); c is a multiple of 4 below d, and the row
// holds at least the next multiple of 4 floats, so the load stays inside it.
__device__ __forceinline__ float4 load_coords4(const float *row, uint32_t c, uint32_t d) {
  float4 v = *reinterpret_cast<const float4 *>(row + c);
  if (c + 1 >= d) v.y = 0.f;
  if (c + 2 >= d) v.z = 0.f;
  if (c + 3 >= d) v.w = 0.f;
  return v;
}

template <int METRIC, bool VQ>
__global__ __launch_bounds__(256) void qr_dist(const float *Q, size_t ldq, uint32_t m, const float *Rp, size_t ldr,
                                               uint32_t R, uint32_t d, float *D) {
  __shared__ __attribute__((aligned(16))) float Qs[kQRC][kQRP];
  __shared__ __attribute__((aligned(16))) float Rs[kQRC][kQRP];
  const int t = threadIdx.x, ty = t / 16, tx = t % 16;
  const uint32_t q0 = blockIdx.x * kQRT, r0 = blockIdx.y * kQRT;
  float acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = 0.f;
  const uint32_t srow = t / 8, scol = (t % 8) * 4;  // staging: rows srow + 32 s, coordinates scol..+3
  for (uint32_t c0 = 0; c0 < d; c0 += kQRC) {
    __syncthreads();
    const uint32_t c = c0 + scol;
#pragma unroll
    for (int s = 0; s < kQRT / 32; ++s) {
      const uint32_t row = srow + 32 * s;
      float4 qv = make_float4(0.f, 0.f, 0.f, 0.f), rv = qv;
      if (q0 + row < m && c < d) {
        const float *qr = Q + size_t(q0 + row) * ldq;
        if constexpr (VQ) {
          qv = load_coords4(qr, c, d);
        } else {
          qv.x = qr[c];
          qv.y = c + 1 < d ? qr[c + 1] : 0.f;
          qv.z = c + 2 < d ? qr[c + 2] : 0.f;
          qv.w = c + 3 < d ? qr[c + 3] : 0.f;
        }
      }
      if (r0 + row < R && c < d) rv = load_coords4(Rp + size_t(r0 + row) * ldr, c, d);
      Qs[scol][row] = qv.x;
      Qs[scol + 1][row] = qv.y;
      Qs[scol + 2][row] = qv.z;
      Qs[scol + 3][row] = qv.w;
      Rs[scol][row] = rv.x;
      Rs[scol + 1][row] = rv.y;
      Rs[scol + 2][row] = rv.z;
      Rs[scol + 3][row] = rv.w;
    }
    __syncthreads();
    const uint32_t cn = min((uint32_t)kQRC, d - c0);
    for (uint32_t c = 0; c < cn; ++c) {
      const float4 qa = *reinterpret_cast<const float4 *>(&Qs[c][ty * 4]);
      const float4 qb = *reinterpret_cast<const float4 *>(&Qs[c][64 + ty * 4]);
      const float4 ra = *reinterpret_cast<const float4 *>(&Rs[c][tx * 4]);
      const float4 rb = *reinterpret_cast<const float4 *>(&Rs[c][64 + tx * 4]);
      const float qv[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
      const float rv[8] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = fma_term<METRIC>(rv[j], qv[i], acc[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t q = q0 + (i < 4 ? ty * 4 + i : 64 + ty * 4 + i - 4);
    if (q >= m) continue;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t r = r0 + (j < 4 ? tx * 4 + j : 64 + tx * 4 + j - 4);
      if (r < R) D[size_t(q) * R + r] = acc[i][j];
    }
  }
}

// ---- 2. row stats ----------------------------------------------------------------
// One wave per query over its row of R distances: nearest rep (lowest index on ties)
// and an upper bound on the k-th smallest rep distance (k <= 64); optionally the reps of
// the P <= 64 smallest lane minima, the lists the BF16 probe scans.
__global__ __launch_bounds__(256) void row_stats(const float *D, uint32_t m, uint32_t R, uint32_t k,
                                                 uint32_t *near, float *kth, uint32_t P, uint32_t *nearP,
                                                 uint32_t P2, uint32_t *planeR) {
  const uint32_t q = blockIdx.x * 4 + threadIdx.x / kWave;
  if (q >= m) return;
  const int l = lane_id();
  const float *row = D + size_t(q) * R;
  // Lane l scans columns l, l+64, ... in ascending order and keeps its smallest value with
  // the lowest column among equals.
  float lm = FLT_MAX;
  uint32_t li = kDummy;
#pragma unroll 8
  for (uint32_t j = l; j < R; j += kWave) {
    const float v = row[j];
    if (v < lm) {
      lm = v;
      li = j;
    }
  }
  // Sorted by (value, column): lane 0 holds the nearest rep (lowest index on ties); the
  // k-th lane minimum is the k-th smallest of k distinct rep distances, so it bounds the
  // k-th smallest rep distance from above.
  wave_bitonic_sort(lm, li);
  if (l == 0) near[q] = li;
  if (l == (int)k - 1) kth[q] = li != kDummy ? lm : FLT_MAX;
  // The P smallest lane minima: P distinct reps near the query (the nearest among them).
  if (nearP && l < (int)P) nearP[size_t(q) * P + l] = li;
  // The P2 smallest: the reps whose Voronoi planes bound the lists in the plan.
  if (planeR && l < (int)P2) planeR[size_t(q) * P2 + l] = li;
}

// ---- 3b. bounds from the probe ----------------------------------------------------
// gamma_hi: upper bound on the true k-th NN distance value (squared for L2), from the
// k-th smallest FP32 value over k distinct points (the reps, or the probed lists' best);
// tau: candidate cut-off on the computed FP32 values; gam: gamma_hi as a metric distance
// (sqrt for L2) for the pruning tests; cert bit s: probe slot s kept every point of its
// list that could pass tau (bit set also when the slot is unused).
struct BoundArgs {
  const float *pd0, *pd1;        // probe values per slot (ascending), [m][Kp]
  uint32_t Kp, k, m, R;
  const float *rep_kth;          // upper bound on the k-th smallest FP32 rep distance (FLT_MAX if R < k)
  const uint32_t *near, *key1, *list_len;
  double e_probe, e_rep;         // relative error of the probe values and of rep_kth
  const float *qerr;             // per query: absolute error of the rep distances (NULL: e_rep)
  int metric;
  float *tau;
  double *gam;
  uint32_t *cert;
  bool no_cert;                  // every list is scanned again (BF16 multi-list probe)
};

__global__ void probe_bounds(BoundArgs b) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.m) return;
  const float *a0 = b.pd0 + size_t(q) * b.Kp;
  const bool two = b.pd1 != nullptr && b.key1[q] < b.R;
  const float *a1 = two ? b.pd1 + size_t(q) * b.Kp : nullptr;
  // k-th smallest of the union of the two ascending slot lists.
  uint32_t i0 = 0, i1 = 0;
  float kth = FLT_MAX;
  for (uint32_t t = 0; t < b.k; ++t) {
    float va = i0 < b.Kp ? a0[i0] : FLT_MAX;
    float vb = (two && i1 < b.Kp) ? a1[i1] : FLT_MAX;
    if (va <= vb) {
      kth = va;
      ++i0;
    } else {
      kth = vb;
      ++i1;
    }
  }
  double gp = kth >= FLT_MAX ? INFINITY : (double)kth / (1.0 - b.e_probe);
  double gr = INFINITY;
  if (b.rep_kth[q] < FLT_MAX)
    gr = b.qerr ? (double)b.rep_kth[q] + (double)b.qerr[q] : (double)b.rep_kth[q] / (1.0 - b.e_rep);
  double ghi = fmin(gp, gr) * (1.0 + 1e-12);
  double t = ghi * (1.0 + b.e_rep) * (1.0 + 1e-12);
  float tf = t >= (double)FLT_MAX ? INFINITY : (float)t;
  if ((double)tf < t) tf = nextafterf(tf, INFINITY);
  b.tau[q] = tf;
  const double G = b.metric == 0 ? sqrt(ghi) * (1.0 + 1e-12) : ghi;
  b.gam[q] = G;
  if (b.no_cert) {
    b.cert[q] = 0u;
    return;
  }
  // FP32 probe: a list is complete when it is no longer than K' or its K'-th kept value
  // already exceeds tau (every point left out is farther than the bound).
  uint32_t c0 = (b.list_len[b.near[q]] <= b.Kp || a0[b.Kp - 1] > tf) ? 1u : 0u, c1 = 1u;
  if (two) c1 = (b.list_len[b.key1[q]] <= b.Kp || a1[b.Kp - 1] > tf) ? 1u : 0u;
  b.cert[q] = c0 | (c1 << 1);
}

// Second-probe key: the proposal's list when it differs from the nearest rep's, else R
// (no second probe).
__global__ void second_probe_key(const uint32_t *prop, const uint32_t *near, uint32_t m, uint32_t R,
                                 uint32_t *key1) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  uint32_t p = prop[q];
  key1[q] = (p < R && p != near[q]) ? p : R;
}

__global__ void fill_u64(unsigned long long *p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// out[0] = max count, out[1] = sum of counts, out[2] = number of counts above cap.
__global__ __launch_bounds__(256) void cand_stats(const uint32_t *cnt, uint32_t m, uint32_t cap,
                                                  unsigned long long *out) {
  uint32_t mx = 0, over = 0;
  unsigned long long sum = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const uint32_t v = cnt[i];
    mx = max(mx, v);
    sum += v;
    over += v > cap;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, kWave));
    sum += __shfl_xor(sum, o, kWave);
    over += (uint32_t)__shfl_xor((int)over, o, kWave);
  }
  // The block's 4 waves combine in LDS, so that one thread per block touches the three
  // shared counters.
  __shared__ unsigned long long part[4][3];
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    part[w][0] = mx;
    part[w][1] = sum;
    part[w][2] = over;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long bm = 0, bs = 0, bo = 0;
    for (int i = 0; i < 4; ++i) {
      bm = max(bm, part[i][0]);
      bs += part[i][1];
      bo += part[i][2];
    }
    atomicMax(&out[0], bm);
    atomicAdd(&out[1], bs);
    atomicAdd(&out[2], bo);
  }
}

// Multi-list probe: a query whose proposal reached the mass threshold within its P lists
// (count c > 0) probes the network's c best lists instead of its geometric set; the other
// slots get key R (no list, dropped by the counting sort).
// A query probes at least minc of the network's lists, so that its probe entries can hold
// its k candidates.
__global__ void proposal_lists(const uint32_t *plists, const uint32_t *pcnt, uint32_t m, uint32_t P, uint32_t R,
                               uint32_t minc, uint32_t *keys) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  uint32_t c = pcnt[q];
  if (c == 0) return;
  c = min(max(c, minc), P);
  for (uint32_t s = 0; s < P; ++s) {
    const uint32_t j = plists[size_t(q) * P + s];
    keys[size_t(q) * P + s] = s < c && j < R ? j : R;
  }
}

// Multi-list probe, per query (one wave): the P lists' Ke best approximate pairs each are
// merged by approximate value, and the K' best of them are re-ranked with tree-ordered FP32
// distances into (pd, pi) ascending -- K' distinct database points, so the k-th of them
// bounds the k-th NN distance from above.
template <int METRIC, int NPL>
__global__ __launch_bounds__(256) void probe_merge(const float *Q, size_t ldq, const float *xp, size_t ldx,
                                                   const uint32_t *inv, uint32_t m, uint32_t d, uint32_t P,
                                                   uint32_t Ke, uint32_t Kp, const float *pv,
                                                   const uint32_t *pid, uint32_t *pi, float *pd, uint32_t n) {
  const uint32_t q = blockIdx.x * 4 + threadIdx.x / kWave;
  if (q >= m) return;
  const int l = lane_id();
  const float *qrow = Q + size_t(q) * ldq;
  float qv[NPL];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    uint32_t c = l + kWave * j;
    qv[j] = c < d ? qrow[c] : 0.f;
  }
  float tv = FLT_MAX;
  uint32_t ti = kDummy;
  const uint32_t total = P * Ke;
  const float *vq = pv + size_t(q) * total;
  const uint32_t *iq = pid + size_t(q) * total;
  // Entries of 16 arrive sorted by (value, slot): the lanes of every other 16-run read
  // theirs backwards, and the 32- and 64-wide stages finish the sort.  Equal values may
  // be out of id order, which only changes which of equally close points are re-ranked.
  const bool runs = Ke == 16;
  for (uint32_t b0 = 0; b0 < total; b0 += kWave) {
    const uint32_t e = runs && (l & 16) ? b0 + (l & ~15) + (15 - (l & 15)) : b0 + l;
    float v = e < total ? vq[e] : FLT_MAX;
    uint32_t id = e < total ? iq[e] : kDummy;
    if (id == kDummy) v = FLT_MAX;
    if (runs)
      wave_bitonic_sort_from16(v, id);
    else
      wave_bitonic_sort(v, id);
    wave_merge_lower(tv, ti, v, id);
  }
  // Lane l < K' holds the l-th best; the re-rank gathers their rows eight at a time.
  const uint32_t myid = l < (int)Kp && ti != kDummy ? RBC_CHECKED(ti, n) : kDummy;
  float rv = FLT_MAX;
  uint32_t ri = kDummy;
  rerank_candidates<METRIC, NPL>(rv, ri, Kp, [&](uint32_t) { return myid; }, xp, ldx, inv, qv, d);
  if (l < (int)Kp) {
    pi[size_t(q) * Kp + l] = ri;
    pd[size_t(q) * Kp + l] = rv;
  }
}

// ---- 4. plan ------------------------------------------------------------------------
// Pass A: keep bits, list-major (R x Wq words, bit b of word (j, t) = query 32 t + b
// keeps list j), and per-list query counts.  Block (x, y) = 4 waves over lists 64y..
// (lane = list) and queries kPlanCQ x.. (wave w takes queries 16w.. of them, and waves
// 2h, 2h+1 make up word h).  Per (query, list) pair every load is independent of the
// others; the per-query values are wave-uniform.
constexpr uint32_t kPlanCQ = 64;  // plan_count: queries per block (fewer list_cnt atomics)

// Lane l of a block covers the V adjacent lists jb + V l .. + V - 1 (V = 4 when rows of
// R floats keep 16-byte alignment), so that every distance and plane-denominator load
// moves V values: the kernel is bound by the number of load instructions.
template <int V>
__device__ __forceinline__ void load_v(const float *p, float (&o)[V]) {
  if constexpr (V == 4) {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    o[0] = t.x, o[1] = t.y, o[2] = t.z, o[3] = t.w;
  } else {
    o[0] = p[0];
  }
}

// The largest float <= v (v not NaN), without branches.
__device__ __forceinline__ float f32_round_down(double v) {
  const float f = (float)v;
  const int b = __float_as_int(f);
  const float below = f > 0.f ? __int_as_float(b - 1) : (f < 0.f ? __int_as_float(b + 1) : -__int_as_float(1));
  return (double)f > v ? below : f;
}

template <int NPR, int V>
__global__ __launch_bounds__(256) void plan_count(const float *Dqr, uint32_t m, uint32_t R,
                                                  const uint32_t *near,
                                                  const uint32_t *key1, const double *gam,
                                                  const uint32_t *cert, const float *radius,
                                                  const float *rr_den, double e, const float *qerr, int metric,
                                                  const uint32_t *planeR, uint32_t P,
                                                  uint32_t Wq, uint32_t *bits, uint32_t *list_cnt) {
  static_assert(kPlanCQ == 64, "four waves of 16 queries make two 32-query words");
  __shared__ uint32_t red[4][kWave][V];
  const int l = lane_id(), w = threadIdx.x / kWave;
  const uint32_t jl = blockIdx.y * (kWave * V) + V * l;  // first list of this lane
  const bool live = jl < R;  // V divides R when V > 1
  // Per list: psi, and the owner slack 2 e psi^2 rounded up; (1 - e) rounded down.
  double psi[V];
  float psf[V], slu[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    psf[v] = live ? radius[jl + v] : 0.f;
    psi[v] = (double)psf[v];
    slu[v] = -f32_round_down(-(2.0 * e * psi[v] * psi[v]));
  }
  const float c1m = f32_round_down(1.0 - e);
  uint32_t mask[V];
#pragma unroll
  for (int v = 0; v < V; ++v) mask[v] = 0;
  for (uint32_t qi = 0; qi < 16; ++qi) {
    const uint32_t q = blockIdx.x * kPlanCQ + 16 * w + qi;
    if (q >= m) break;
    const uint32_t j0 = near[q];
    const uint32_t j1 = key1 ? key1[q] : R;
    const uint32_t cb = cert[q];
    const double Eq = qerr ? (double)qerr[q] : 0.0;  // wave-uniform, loaded with the others
    const double G = gam[q];
    if (!live) continue;
    float dq[V];
    load_v<V>(Dqr + size_t(q) * R + jl, dq);
    bool keep[V], tri = false;
    float A[V];
    const double Gp = G * (1.0 + 2e-12);
    const double Gd = Gp * (1.0 + 1e-13);
    // The probed lists j0 and j1 are rescanned only if their probe was not complete.
    if (metric == 0) {
      // list_lower_bound's test without square roots or divisions: the triangle bound
      // sqrt(Dqj (1-e)) - psi <= G' (G' = G / (1 - 1e-12)) as Dqj (1-e) <= (G' + psi)^2,
      // and each plane bound num / den <= G' as num <= G' den with den rounded up; both
      // keep at least what it keeps.  With split-BF16 rep distances, D~ is within qerr[q]
      // of the true value (absolute), and the lower value Dlo and the upper values B_i
      // below carry that term too.  In FP32 with every rounding towards keeping: Dlo =
      // fma(Dqj, (1-e) rounded down, -Eq) less 2^-22 of its magnitude (an FMA is within
      // 2^-24 of the exact value), (G' + psi)^2 from G' rounded up, times (1 + 2^-20)
      // (three roundings of at most 2^-24 each, and the 1e-13).  Branch-free over the
      // lane's lists.
      const float Eqf = qerr ? qerr[q] : 0.f;
      const float Gpf = -f32_round_down(-Gp);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const uint32_t j = jl + v;
        const float x = __builtin_fmaf(dq[v], c1m, -Eqf);
        const float dlo = x - fabsf(x) * 0x1p-22f;
        const float rhs = Gpf + psf[v];
        const bool kt = dlo <= (rhs * rhs) * (1.f + 0x1p-20f);
        keep[v] = j == j0 ? (cb & 1u) == 0 : (j == j1 ? (cb & 2u) == 0 : kt);
        tri |= kt & (j != j0) & (j != j1);
        // The plane tests' left side Dlo - slack, rounded down the same way.
        const float a = dlo - slu[v];
        A[v] = a - fabsf(a) * 0x1p-22f;
      }
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const uint32_t j = jl + v;
        A[v] = 0.f;
        const double lb = list_lower_bound((double)dq[v], 0.0, 0.0, psi[v], e, true, metric);
        keep[v] = j == j0 ? (cb & 1u) == 0 : (j == j1 ? (cb & 2u) == 0 : lb <= G);
      }
    }
    // The planes only matter for lists the triangle bound keeps; their rr_den segments are
    // not read when none of the lane's lists was kept by it.
    if (tri) {
      // The plane between r_j and r_i: an x owned by r_j satisfies |x-r_j|^2 <= |x-r_i|^2 +
      // 2e psi^2 for EVERY rep r_i (its owner is the FP32 argmin over all reps), so the
      // nearest rep and each plane rep give a lower bound, and the largest decides.  All
      // loads first (independent of each other), then the tests.
      uint32_t ji[1 + NPR];
      float Dqi[1 + NPR], den[1 + NPR][V];
      ji[0] = j0;
#pragma unroll
      for (uint32_t pi = 0; pi < (uint32_t)NPR; ++pi) {
        const uint32_t r = pi < P ? planeR[size_t(q) * P + pi] : R;
        ji[1 + pi] = r < R && r != j0 ? r : R;
      }
#pragma unroll
      for (uint32_t pi = 0; pi <= (uint32_t)NPR; ++pi) {
        const bool ok = ji[pi] < R;
        Dqi[pi] = ok ? Dqr[size_t(q) * R + ji[pi]] : 0.f;
        if (ok) {
          load_v<V>(rr_den + size_t(ji[pi]) * R + jl, den[pi]);
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) den[pi][v] = 0.f;
        }
      }
      // The plane test num > G' den in FP32, with every rounding pushed towards keeping:
      // A <= Dlo - slack (above), and fl(Gf den + B_i) >= G' den + Dqi (1+e) + Eq,
      // since Gf >= G' (1 + 8u) and B_i = fl(Dqi fl(1 + e + 8u) + Eu) >= (Dqi (1+e) + Eq)
      // (1 + 6u) with Eu >= Eq (1 + 8u) (u = 2^-24).  A > that implies the exact test.
      // A rep's plane against its own list (ji == j) is skipped (no bound).
      const double Eud = Eq * (1.0 + 0x1p-21);
      float Eu = (float)Eud;
      if ((double)Eu < Eud) Eu = nextafterf(Eu, INFINITY);
      const double Gu = Gd * (1.0 + 0x1p-21);
      float Gf = (float)Gu;
      if ((double)Gf < Gu) Gf = nextafterf(Gf, INFINITY);
      const float cb1 = (float)(1.0 + e + 0x1p-21);
      // B_i depends on the plane rep only; a missing plane rep has den = 0, and a plane
      // with den <= 0 or against the rep's own list never prunes.  Branch-free over the
      // lane's lists: the probed lists j0 and j1 and those the triangle dropped keep
      // their decision.
      float Bi[1 + NPR];
#pragma unroll
      for (uint32_t pi = 0; pi <= (uint32_t)NPR; ++pi) Bi[pi] = __builtin_fmaf(Dqi[pi], cb1, Eu);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        bool far = false;
#pragma unroll
        for (uint32_t pi = 0; pi <= (uint32_t)NPR; ++pi) {
          const float dn = den[pi][v];
          far |= (dn > 0.f) & (ji[pi] != jl + v) & (A[v] > __builtin_fmaf(Gf, dn, Bi[pi]));
        }
        keep[v] = keep[v] & !(far & (jl + v != j0) & (jl + v != j1));
      }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) mask[v] |= uint32_t(keep[v]) << qi;
  }
#pragma unroll
  for (int v = 0; v < V; ++v) red[w][l][v] = mask[v];
  __syncthreads();
  if (w == 0 && live) {
    const uint32_t t0 = blockIdx.x * 2;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const uint32_t j = jl + v;
      const uint32_t w0 = red[0][l][v] | (red[1][l][v] << 16), w1 = red[2][l][v] | (red[3][l][v] << 16);
      if (t0 < Wq) bits[size_t(j) * Wq + t0] = w0;
      if (t0 + 1 < Wq) bits[size_t(j) * Wq + t0 + 1] = w1;
      const uint32_t tot = __popc(w0) + __popc(w1);
      if (tot) atomicAdd(&list_cnt[j], tot);
    }
  }
}

// Pass C: scatter query ids list-major.  One wave per (list blockIdx.y, 64 words = 2048
// queries): lane l reads word w0 + l of the list's row (one 256-byte segment), a wave
// prefix sum of the popcounts places every lane's ids after those of the lanes before it,
// and one atomic per wave reserves the wave's range of the list.  The ids are placed in
// the wave's LDS buffer first and then copied out with coalesced stores.
__global__ __launch_bounds__(256) void plan_scatter(uint32_t Wq, const uint32_t *bits,
                                                    const uint32_t *list_qoff, uint32_t *cursor,
                                                    uint32_t *qlist, uint32_t cap) {
  __shared__ uint32_t stage[4][32 * kWave];
  const uint32_t j = blockIdx.y;
  const int wv = threadIdx.x / kWave;
  const uint32_t w0 = (blockIdx.x * 4 + wv) * kWave;
  if (w0 >= Wq) return;
  const int l = lane_id();
  const uint32_t t = w0 + l;
  uint32_t word = t < Wq ? bits[size_t(j) * Wq + t] : 0u;
  const uint32_t c = __popc(word);
  uint32_t incl = c;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o, kWave);
    if (l >= o) incl += v;
  }
  const uint32_t total = __shfl(incl, kWave - 1, kWave);
  if (total == 0) return;
  uint32_t base = 0;
  if (l == 0) base = list_qoff[j] + atomicAdd(&cursor[j], total);
  base = __shfl(base, 0, kWave);
  uint32_t pos = incl - c;
  while (word) {
    stage[wv][pos++] = t * 32 + (__ffs(word) - 1);
    word &= word - 1;
  }
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = l; i < total; i += kWave) qlist[RBC_CHECKED(base + i, cap)] = stage[wv][i];
}

}  // namespace

// ---- 6. select ----------------------------------------------------------------------

template <int METRIC, int NPL>
__global__ __launch_bounds__(256) void select_kernel(SelectArgs a) {
  const uint32_t q = blockIdx.x * 4 + threadIdx.x / kWave;
  if (q >= a.m) return;
  const int l = lane_id();
  const float *qrow = a.q + size_t(q) * a.ldq;
  float qv[NPL];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    uint32_t c = l + kWave * j;
    qv[j] = c < a.d ? qrow[c] : 0.f;
  }
  float tv = FLT_MAX;
  uint32_t ti = kDummy;
  const uint32_t cb = a.cert[q];
  const uint32_t nc = a.cand_n ? min(a.cand_n[q], a.cap) : 0;
  if (a.probe_d) {
    // Start from the probe's exactly re-ranked points (sorted, lanes < Kp), and re-rank
    // only the scan candidates not among them, 64 at a time, compacted through LDS.
    __shared__ uint32_t comp[4][kWave];
    const int w = threadIdx.x / kWave;
    uint32_t pid = kDummy;
    if (l < (int)a.Kp) {
      pid = a.probe_i[size_t(q) * a.Kp + l];
      if (pid != kDummy) {
        ti = RBC_CHECKED(pid, a.n);
        tv = a.probe_d[size_t(q) * a.Kp + l];
      }
    }
    uint32_t seen = (uint32_t)__popcll(__ballot(pid != kDummy));
    for (uint32_t b0 = 0; b0 < nc; b0 += kWave) {
      uint32_t id = b0 + l < nc ? a.cand_i[size_t(q) * a.cap + b0 + l] : kDummy;
      for (uint32_t j = 0; j < a.Kp; ++j)
        if (id == (uint32_t)__shfl((int)pid, (int)j, kWave)) id = kDummy;
      const uint64_t bal = __ballot(id != kDummy);
      const uint32_t nv = (uint32_t)__popcll(bal);
      if (id != kDummy) comp[w][__popcll(bal & ((1ull << l) - 1ull))] = id;
      __builtin_amdgcn_wave_barrier();
      auto cand = [&](uint32_t e) -> uint32_t {
        const uint32_t id = comp[w][e];
        return RBC_CHECKED(id, a.n);
      };
      rerank_candidates<METRIC, NPL>(tv, ti, nv, cand, a.xp, a.ldx, a.inv, qv, a.d);
      __builtin_amdgcn_wave_barrier();
      seen += nv;
    }
    near_tie_guard<METRIC>(tv, ti, seen, a.xp, a.ldx, a.inv, qrow, a.d, a.k, a.e_tree, a.stats);
  } else {
    const uint32_t np0 = (cb & 1u) ? a.Kp : 0;
    const bool two = a.probe_i1 != nullptr && a.key1[q] < a.R && (cb & 2u);
    const uint32_t np = np0 + (two ? a.Kp : 0);
    const uint32_t total = np + nc;
    auto cand = [&](uint32_t e) -> uint32_t {
      uint32_t id;
      if (e < np0) id = a.probe_i[size_t(q) * a.Kp + e];
      else if (e < np) id = a.probe_i1[size_t(q) * a.Kp + (e - np0)];
      else id = a.cand_i[size_t(q) * a.cap + (e - np)];
      return id == kDummy ? id : RBC_CHECKED(id, a.n);
    };
    rerank_candidates<METRIC, NPL>(tv, ti, total, cand, a.xp, a.ldx, a.inv, qv, a.d);
    near_tie_guard<METRIC>(tv, ti, total, a.xp, a.ldx, a.inv, qrow, a.d, a.k, a.e_tree, a.stats);
  }
  if (l < (int)a.k) {
    a.out_d[size_t(q) * a.ldo + l] = tv;
    a.out_i[size_t(q) * a.ldo + l] = ti;
  }
}

// Exact FP32 (tree-order) distances of a probe's K' candidates, re-sorted by
// (distance, id); one wave per query.

int launch_probe_merge(const float *Q, size_t ldq, const rbc_exact *h, uint32_t m, uint32_t P, uint32_t Ke,
                       uint32_t Kp, const float *pv, const uint32_t *pid, uint32_t *pi, float *pd,
                       hipStream_t st) {
  uint32_t T = tree_width(h->d);
  dim3 g(cdiv(m, 4)), b(256);
  if (T <= 64)
    hipLaunchKernelGGL((probe_merge<0, 1>), g, b, 0, st, Q, ldq, h->xp.get(), (size_t)h->ld, h->inv.get(), m,
                       h->d, P, Ke, Kp, pv, pid, pi, pd, h->n);
  else if (T == 128)
    hipLaunchKernelGGL((probe_merge<0, 2>), g, b, 0, st, Q, ldq, h->xp.get(), (size_t)h->ld, h->inv.get(), m,
                       h->d, P, Ke, Kp, pv, pid, pi, pd, h->n);
  else
    hipLaunchKernelGGL((probe_merge<0, 4>), g, b, 0, st, Q, ldq, h->xp.get(), (size_t)h->ld, h->inv.get(), m,
                       h->d, P, Ke, Kp, pv, pid, pi, pd, h->n);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int launch_select(const SelectArgs &a, int metric, hipStream_t st) {
  uint32_t P = tree_width(a.d);
  dim3 g(cdiv(a.m, 4)), b(256);
#define RBC_SEL(M, N) hipLaunchKernelGGL((select_kernel<M, N>), g, b, 0, st, a)
  if (metric == 0) {
    if (P <= 64) RBC_SEL(0, 1);
    else if (P == 128) RBC_SEL(0, 2);
    else RBC_SEL(0, 4);
  } else {
    if (P <= 64) RBC_SEL(1, 1);
    else if (P == 128) RBC_SEL(1, 2);
    else RBC_SEL(1, 4);
  }
#undef RBC_SEL
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

// ---- host orchestration ---------------------------------------------------------------

// Workspace slots of the query.  Each primitive call gets its own window of kPrimSlots
// slots from 0: an exclusive scan takes 2 slots per recursion level (at most 3 levels
// below 2^33 elements), a counting sort or an item generation one more, so a window never
// reaches the next one, and the query's own buffers start at kQ above all windows.
constexpr size_t kPrimSlots = 10;
constexpr size_t kSlotSort = 0, kSlotProbeItems = 10, kSlotPlanScan = 20, kSlotScanItems = 30, kQ = 64;
static_assert(kPrimSlots >= 1 + 2 * 3 && kSlotScanItems + kPrimSlots <= kQ, "workspace slot windows overlap");

// Creates the side stream and its two events together (all or none).
int ensure_side_stream(rbc_exact *h) {
  if (h->side && h->ev_in && h->ev_out) return RBC_OK;
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
  if (e != hipSuccess) {
    if (b) (void)hipEventDestroy(b);
    if (a) (void)hipEventDestroy(a);
    if (s) (void)hipStreamDestroy(s);
    set_error("creating the proposal's side stream failed: %s", hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RBC_ERR_NOMEM : RBC_ERR_HIP;
  }
  h->side = s;
  h->ev_in = a;
  h->ev_out = b;
  return RBC_OK;
}

// Waits for the side stream when the query returns (its work is done by then on success;
// on an error return this keeps it from outliving the call).
struct SideStreamGuard {
  hipStream_t s = nullptr;
  ~SideStreamGuard() {
    if (s) (void)hipStreamSynchronize(s);
  }
};

struct Timer {
  hipEvent_t ev[8];
  int n = 0;
  bool on = false;
  hipStream_t st;
  void init(bool enable, hipStream_t s) {
    on = enable;
    st = s;
    if (on)
      for (auto &e : ev) (void)hipEventCreate(&e);
  }
  void mark() {
    if (on && n < 8) (void)hipEventRecord(ev[n++], st);
  }
  float ms(int a, int b) {
    float t = 0.f;
    if (on && b < n) (void)hipEventElapsedTime(&t, ev[a], ev[b]);
    return t;
  }
  ~Timer() {
    if (on)
      for (auto &e : ev) (void)hipEventDestroy(e);
  }
};

// Step 1 (q -> reps) into D (m x R).  BF16 filter: also the queries' BF16 copies (qb,
// qn, qe) and, for L2, the split-BF16 MFMA distances with their per-query absolute error
// bounds (qn2, qerr); otherwise FP32 VALU distances (relative error fma_rel_err(dpad)).
static int rep_distances(rbc_exact *h, const float *qd, size_t ldq, uint32_t m, int precision, float *Dqr,
                         uint16_t *qb, float *qn, float *qe, float *qn2, float *qerr, uint16_t *qlo, hipStream_t st) {
  const uint32_t R = h->R, d = h->d;
  const bool qr_split = precision == RBC_PREC_BF16 && h->metric == RBC_L2;
  if (precision == RBC_PREC_BF16) {
    RBC_TRY(bf16_prepare_queries(h, qd, ldq, m, qb, qn, qe, st, qn2, qerr, qr_split ? qlo : nullptr));
    if (qr_split) RBC_TRY(bf16_qr_dist(h, qb, qlo, m, qn2, Dqr, st));
  }
  if (!qr_split) {
    const dim3 g(cdiv(m, kQRT), cdiv(R, kQRT));
    // 16-byte query loads need 16-byte aligned rows (the reps are padded to 16 floats).
    const bool vq = ldq % 4 == 0 && reinterpret_cast<uintptr_t>(qd) % 16 == 0;
    const float *rp = h->reps.get();
    const size_t ldr = h->ld;
    if (h->metric == RBC_L2) {
      if (vq)
        hipLaunchKernelGGL((qr_dist<0, true>), g, dim3(256), 0, st, qd, ldq, m, rp, ldr, R, d, Dqr);
      else
        hipLaunchKernelGGL((qr_dist<0, false>), g, dim3(256), 0, st, qd, ldq, m, rp, ldr, R, d, Dqr);
    } else {
      if (vq)
        hipLaunchKernelGGL((qr_dist<1, true>), g, dim3(256), 0, st, qd, ldq, m, rp, ldr, R, d, Dqr);
      else
        hipLaunchKernelGGL((qr_dist<1, false>), g, dim3(256), 0, st, qd, ldq, m, rp, ldr, R, d, Dqr);
    }
    RBC_HIP_TRY(hipGetLastError());
  }
  return RBC_OK;
}

static int exact_query_run(rbc_exact *h, const float *qd, size_t ldq, uint32_t m, uint32_t k, int precision,
                       float *od, uint32_t *oi, uint32_t ldo, hipStream_t st, rbc_query_stats *stats) {
  const uint32_t R = h->R, d = h->d;
  Workspace &ws = h->ws;
  Timer tm;
  tm.init(stats != nullptr, st);
  tm.mark();
  const double e = fma_rel_err(h->dpad);
  const uint32_t Kp = std::min<uint32_t>(RBC_MAX_K, k + 8);
  void *p;
  // 2'. The proposal network's inference needs only the queries: it runs on the side
  // stream while the main stream computes the q -> reps distances, and proposal_slot
  // waits for it.
  uint32_t *plists = nullptr, *pcnt = nullptr;
  // An early error return must not leave the proposal's inference running on the side
  // stream over workspace that a later call reuses.
  SideStreamGuard side_guard;
  const uint32_t Pprop = std::min<uint32_t>(h->probe_lists, R);
  if (h->proposal && precision == RBC_PREC_BF16 && k <= 32) {
    RBC_TRY(ws.get(kQ + 23, size_t(m) * Pprop * 4, &p));
    plists = static_cast<uint32_t *>(p);
    RBC_TRY(ws.get(kQ + 37, size_t(m) * 4, &p));
    pcnt = static_cast<uint32_t *>(p);
    RBC_TRY(ensure_side_stream(h));
    RBC_HIP_TRY(hipEventRecord(h->ev_in, st));
    RBC_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_in, 0));
    side_guard.s = h->side;
    RBC_TRY(mlp_propose_lists(h->proposal, qd, ldq, m, Pprop, h->prop_mass, plists, pcnt, h->side));
    RBC_HIP_TRY(hipEventRecord(h->ev_out, h->side));
  }
  // 1. q -> reps.  BF16 filter, L2: split-BF16 MFMA distances with a per-query absolute
  // error bound (qerr); otherwise FP32 VALU distances (relative error e).
  RBC_TRY(ws.get(kQ + 0, size_t(m) * R * 4, &p));
  float *Dqr = static_cast<float *>(p);
  uint16_t *qb = nullptr;
  float *qn = nullptr, *qe = nullptr, *qerr = nullptr;
  if (precision == RBC_PREC_BF16) {
    RBC_TRY(ws.get(kQ + 20, size_t(m) * h->ldb * 2, &p));
    qb = static_cast<uint16_t *>(p);
    RBC_TRY(ws.get(kQ + 21, size_t(m) * 4, &p));
    qn = static_cast<float *>(p);
    RBC_TRY(ws.get(kQ + 22, size_t(m) * 4, &p));
    qe = static_cast<float *>(p);
  }
  float *qn2 = nullptr;
  uint16_t *qlo = nullptr;
  if (precision == RBC_PREC_BF16 && h->metric == RBC_L2) {
    RBC_TRY(ws.get(kQ + 38, size_t(m) * 4, &p));
    qn2 = static_cast<float *>(p);
    RBC_TRY(ws.get(kQ + 39, size_t(m) * 4, &p));
    qerr = static_cast<float *>(p);
    RBC_TRY(ws.get(kQ + 44, size_t(m) * h->ldb * 2, &p));
    qlo = static_cast<uint16_t *>(p);
  }
  RBC_TRY(rep_distances(h, qd, ldq, m, precision, Dqr, qb, qn, qe, qn2, qerr, qlo, st));
  const bool bf16 = precision == RBC_PREC_BF16;
  // BF16 filter on an index whose data lie close to a low-dimensional subspace: the scan
  // tests the queries' and points' projected sketches (exact_sketch.hip).
  const bool sketch = bf16 && h->metric == RBC_L2 && h->sk_active;
  uint16_t *qs = nullptr;
  float *qs_e = nullptr, *qs_n = nullptr;
  if (sketch) {
    RBC_TRY(ws.get(kQ + 40, size_t(m) * kSkW * 2, &p));
    qs = static_cast<uint16_t *>(p);
    RBC_TRY(ws.get(kQ + 41, size_t(m) * 4, &p));
    qs_e = static_cast<float *>(p);
    RBC_TRY(ws.get(kQ + 42, size_t(m) * 4, &p));
    qs_n = static_cast<float *>(p);
    RBC_TRY(sketch_queries(h, qd, ldq, m, qs, qs_e, qs_n, st));
  }
  // The BF16 probe keeps at most 32 candidates, so it serves k <= 32.  It scans the Pp
  // lists of the smallest lane minima of each query's rep distances in one pass.  Its K'
  // best by approximate distance are re-ranked exactly, and gamma is the k-th of them:
  // K' = k + 8 on full BF16 rows, k + 2 on sketches (on low-rank data, n = 4M, d = 128,
  // 16-dimensional, k + 2, 4 and 8 leave the same plan; k + 2 re-ranks 1/3 fewer rows).
  const bool bprobe = bf16 && k <= 32;
  const uint32_t KpUse = bprobe ? std::min<uint32_t>(32, k + (sketch ? 2u : 8u)) : Kp;
  const uint32_t Pp = bprobe ? std::min<uint32_t>(h->probe_lists, R) : 0;
  // 2. row stats
  RBC_TRY(ws.get(kQ + 1, size_t(m) * 4, &p));
  uint32_t *near = static_cast<uint32_t *>(p);
  RBC_TRY(ws.get(kQ + 2, size_t(m) * 4, &p));
  float *rep_kth = static_cast<float *>(p);
  uint32_t *nearP = nullptr, *planeR = nullptr;
  const uint32_t prep = h->plane_reps != rbc_exact::kPlaneRepsAuto ? h->plane_reps : sketch ? 8u : 12u;
  const uint32_t P2 = bprobe ? std::min<uint32_t>(prep, R) : 0;
  if (bprobe) {
    RBC_TRY(ws.get(kQ + 31, size_t(m) * Pp * 4, &p));
    nearP = static_cast<uint32_t *>(p);
    RBC_TRY(ws.get(kQ + 36, size_t(m) * P2 * 4, &p));
    planeR = static_cast<uint32_t *>(p);
  }
  RBC_TRY(fill_f32(rep_kth, FLT_MAX, m, st));
  hipLaunchKernelGGL(row_stats, dim3(cdiv(m, 4)), dim3(256), 0, st, Dqr, m, R, std::min(k, R), near,
                     rep_kth, Pp, nearP, P2, planeR);
  RBC_HIP_TRY(hipGetLastError());
  if (R < k) RBC_TRY(fill_f32(rep_kth, FLT_MAX, m, st));
  tm.mark();
  // 3. probe.  FP32: group queries by nearest rep and scan that list with FP32 tiles
  // keeping the K' best (and the proposal's list likewise); lists whose probe is complete
  // are not scanned again.  BF16: every query's Pp probe lists are scanned in one pass of
  // BF16 MFMA tiles keeping the K' best approximate distances per (query, list); the
  // query's K' best over its lists are re-ranked in FP32.  Only the bound is used from it:
  // the scan visits every surviving list, probed or not.
  DevBuf<ScanItem> pitems;
  uint32_t npitems = 0;
  // Runs one probe slot for the queries whose key (list) is < R.
  auto run_probe = [&](const uint32_t *keys, size_t slot_order, size_t slot_off, size_t slot_d, size_t slot_i,
                       float **pd_out, uint32_t **pi_out) -> int {
    void *pp;
    RBC_TRY(ws.get(slot_order, size_t(m) * 4, &pp));
    uint32_t *order = static_cast<uint32_t *>(pp);
    RBC_TRY(ws.get(slot_off, (size_t(R) + 1) * 4, &pp));
    uint32_t *off = static_cast<uint32_t *>(pp);
    RBC_TRY(counting_sort(keys, m, R, order, off, ws, kSlotSort, st));
    RBC_TRY(make_bucket_items(off, h->list_off.get(), h->list_len.get(), R, pitems, &npitems, ws,
                              kSlotProbeItems, st, (uint32_t)kTQ));
    RBC_TRY(ws.get(slot_d, size_t(m) * KpUse * 4, &pp));
    float *pd = static_cast<float *>(pp);
    RBC_TRY(ws.get(slot_i, size_t(m) * KpUse * 4, &pp));
    uint32_t *pi = static_cast<uint32_t *>(pp);
    RBC_TRY(fill_f32(pd, FLT_MAX, size_t(m) * KpUse, st));
    RBC_TRY(fill_u32(pi, kDummy, size_t(m) * KpUse, st));
    TileSrc src{h->xp.get(), h->ld, nullptr, h->perm.get(), qd, ldq, order, d, h->dpad};
    EpiTopK epi{KpUse, pd, pi, KpUse, 0};
    RBC_TRY(launch_tiles_m(h->metric, src, pitems.get(), npitems, epi, st));
    *pd_out = pd;
    *pi_out = pi;
    return RBC_OK;
  };
  float *probe_d = nullptr, *probe_d1 = nullptr;
  uint32_t *probe_i = nullptr, *probe_i1 = nullptr, *key1 = nullptr;
  if (bprobe) {
    // Entries e = q * Pp + s (query q, probe slot s), keyed by list, sorted list-major.
    // Each entry holds at most Ke candidates (the sketch probe 8, the full-row probe 16);
    // the merge reads that many.  A query taking the network's lists probes enough of them
    // to hold its k (ceil(k / Ke)).
    const uint32_t Ke = std::min(KpUse, sketch ? 8u : kBf16ProbeKeep);
    if (plists) {
      RBC_HIP_TRY(hipStreamWaitEvent(st, h->ev_out, 0));
      hipLaunchKernelGGL(proposal_lists, dim3(cdiv(m, 256)), dim3(256), 0, st, plists, pcnt, m, Pp, R,
                         cdiv(k, Ke), nearP);
      RBC_HIP_TRY(hipGetLastError());
    }
    const size_t ne = size_t(m) * Pp;
    RBC_TRY(ws.get(kQ + 3, ne * 4, &p));
    uint32_t *order = static_cast<uint32_t *>(p);
    RBC_TRY(ws.get(kQ + 4, (size_t(R) + 1) * 4, &p));
    uint32_t *off = static_cast<uint32_t *>(p);
    RBC_TRY(counting_sort(nearP, ne, R, order, off, ws, kSlotSort, st));
    RBC_TRY(make_bucket_items(off, h->list_off.get(), h->list_len.get(), R, pitems, &npitems, ws,
                              kSlotProbeItems, st,
                              kBf16ProbeQ));
    RBC_TRY(ws.get(kQ + 32, ne * Ke * 4, &p));
    float *ev = static_cast<float *>(p);
    RBC_TRY(ws.get(kQ + 33, ne * Ke * 4, &p));
    uint32_t *eid = static_cast<uint32_t *>(p);
    RBC_TRY(fill_f32(ev, FLT_MAX, ne * Ke, st));
    RBC_TRY(fill_u32(eid, kDummy, ne * Ke, st));
    if (sketch)
      RBC_TRY(sketch_probe(h, qs, qs_n, order, Pp, pitems.get(), npitems, Ke, eid, ev, st));
    else
      RBC_TRY(bf16_probe(h, qb, qn, order, Pp, pitems.get(), npitems, Ke, eid, ev, st));
    RBC_TRY(ws.get(kQ + 5, size_t(m) * KpUse * 4, &p));
    probe_d = static_cast<float *>(p);
    RBC_TRY(ws.get(kQ + 6, size_t(m) * KpUse * 4, &p));
    probe_i = static_cast<uint32_t *>(p);
    RBC_TRY(launch_probe_merge(qd, ldq, h, m, Pp, Ke, KpUse, ev, eid, probe_i, probe_d, st));
  } else {
    RBC_TRY(run_probe(near, kQ + 3, kQ + 4, kQ + 5, kQ + 6, &probe_d, &probe_i));
  }
  // 3a. FP32 second probe: the proposal network's top-1 list where it differs from the
  // nearest rep's.
  if (h->proposal && !bprobe) {
    RBC_TRY(ws.get(kQ + 23, size_t(m) * 4, &p));
    uint32_t *prop = static_cast<uint32_t *>(p);
    RBC_TRY(mlp_propose(h->proposal, qd, ldq, m, prop, st));
    RBC_TRY(ws.get(kQ + 24, size_t(m) * 4, &p));
    key1 = static_cast<uint32_t *>(p);
    hipLaunchKernelGGL(second_probe_key, dim3(cdiv(m, 256)), dim3(256), 0, st, prop, near, m, R, key1);
    RBC_HIP_TRY(hipGetLastError());
    RBC_TRY(run_probe(key1, kQ + 25, kQ + 26, kQ + 27, kQ + 28, &probe_d1, &probe_i1));
  }
  // 3b. bounds
  RBC_TRY(ws.get(kQ + 7, size_t(m) * 4, &p));
  float *tau = static_cast<float *>(p);
  RBC_TRY(ws.get(kQ + 8, size_t(m) * 8, &p));
  double *gam = static_cast<double *>(p);
  RBC_TRY(ws.get(kQ + 9, size_t(m) * 4, &p));
  uint32_t *cert = static_cast<uint32_t *>(p);
  {
    BoundArgs ba{};
    ba.pd0 = probe_d;
    ba.pd1 = probe_d1;
    ba.Kp = KpUse;
    ba.k = k;
    ba.m = m;
    ba.R = R;
    ba.rep_kth = rep_kth;
    ba.near = near;
    ba.key1 = key1;
    ba.list_len = h->list_len.get();
    ba.e_probe = bprobe ? tree_rel_err(tree_width(d)) : e;
    ba.e_rep = e;
    ba.qerr = qerr;
    ba.metric = h->metric;
    ba.no_cert = bprobe;
    ba.tau = tau;
    ba.gam = gam;
    ba.cert = cert;
    hipLaunchKernelGGL(probe_bounds, dim3(cdiv(m, 256)), dim3(256), 0, st, ba);
    RBC_HIP_TRY(hipGetLastError());
  }
  tm.mark();
  // 4. plan
  const uint32_t Wq = cdiv(m, 32);
  RBC_TRY(ws.get(kQ + 10, size_t(R) * Wq * 4, &p));
  uint32_t *bits = static_cast<uint32_t *>(p);
  RBC_TRY(ws.get(kQ + 11, (size_t(R) + 1) * 4, &p));
  uint32_t *lcnt = static_cast<uint32_t *>(p);
  RBC_TRY(ws.get(kQ + 12, (size_t(R) + 1) * 4, &p));
  uint32_t *cursor = static_cast<uint32_t *>(p);
  RBC_TRY(fill_u32(lcnt, 0, R + 1, st));
  RBC_TRY(fill_u32(cursor, 0, R + 1, st));
  {
    // Four lists per lane when rows of R floats stay 16-byte aligned.
    const bool v4 = R % 4 == 0;
    const dim3 pg(cdiv(m, kPlanCQ), cdiv(R, kWave * (v4 ? 4 : 1)));
#define RBC_PLAN(NPR, V)                                                                                         \
  hipLaunchKernelGGL((plan_count<NPR, V>), pg, dim3(256), 0, st, Dqr, m, R, near, key1, gam, cert,              \
                     h->radius.get(), h->rr_den.get(), e, qerr, h->metric, planeR, P2, Wq, bits, lcnt)
    if (v4) {
      if (P2 <= 4)
        RBC_PLAN(4, 4);
      else if (P2 <= 8)
        RBC_PLAN(8, 4);
      else
        RBC_PLAN(12, 4);
    } else {
      if (P2 <= 4)
        RBC_PLAN(4, 1);
      else if (P2 <= 8)
        RBC_PLAN(8, 1);
      else
        RBC_PLAN(12, 1);
    }
#undef RBC_PLAN
  }
  RBC_HIP_TRY(hipGetLastError());
  RBC_TRY(exclusive_scan_u32(lcnt, lcnt, R, ws, kSlotPlanScan, st));
  RBC_TRY(ws.get(kQ + 13, 4, &p));
  uint32_t total_pairs = 0;
  RBC_HIP_TRY(hipMemcpyAsync(&total_pairs, lcnt + R, 4, hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  RBC_TRY(ws.get(kQ + 14, size_t(std::max(total_pairs, 1u)) * 4, &p));
  uint32_t *qlist = static_cast<uint32_t *>(p);
  hipLaunchKernelGGL(plan_scatter, dim3(cdiv(Wq, 4 * kWave), R), dim3(256), 0, st, Wq, bits, lcnt, cursor,
                     qlist, total_pairs);
  RBC_HIP_TRY(hipGetLastError());
  DevBuf<ScanItem> sitems;
  uint32_t nsitems = 0;
  const uint32_t scan_tile = sketch ? kSkScanQ : bf16 ? kBf16ScanQ : (uint32_t)kTQ;
  RBC_TRY(make_bucket_items(lcnt, h->list_off.get(), h->list_len.get(), R, sitems, &nsitems, ws,
                            kSlotScanItems, st, scan_tile));
  tm.mark();
  // 5. scan with candidate emission; re-run with a larger buffer on overflow.
  uint32_t &cap_seen = h->cand_cap[bf16 ? 1 : 0];
  uint32_t cap = std::max<uint32_t>({256u, 4 * k, cap_seen});
  RBC_TRY(ws.get(kQ + 16, size_t(m) * 4, &p));
  uint32_t *cand_n = static_cast<uint32_t *>(p);
  float *cand_d = nullptr;
  uint32_t *cand_i = nullptr;
  double fallbacks = 0, rerank_total = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    RBC_TRY(ws.get(kQ + 17, size_t(m) * cap * 4, &p));
    cand_d = static_cast<float *>(p);
    RBC_TRY(ws.get(kQ + 18, size_t(m) * cap * 4, &p));
    cand_i = static_cast<uint32_t *>(p);
    RBC_TRY(fill_u32(cand_n, 0, m, st));
    if (sketch) {
      RBC_TRY(ws.get(kQ + 35, 32 * 4, &p));
      uint32_t *sched = static_cast<uint32_t *>(p);
      RBC_TRY(ws.get(kQ + 43, size_t(m) * 4, &p));
      RBC_TRY(sketch_scan(h, m, qs, qs_e, qs_n, gam, static_cast<uint32_t *>(p), qlist, lcnt, sitems.get(), nsitems,
                          sched, cand_i, cand_n, cap, st));
    } else if (precision == RBC_PREC_BF16) {
      RBC_TRY(ws.get(kQ + 35, 32 * 4, &p));
      RBC_TRY(bf16_scan(h, qb, qn, qe, gam, qlist, lcnt, sitems.get(), nsitems, static_cast<uint32_t *>(p),
                        cand_d, cand_i, cand_n, cap, st));
    } else {
      TileSrc src{h->xp.get(), h->ld, nullptr, h->perm.get(), qd, ldq, qlist, d, h->dpad};
      EpiEmit epi{tau, cand_d, cand_i, cand_n, cap};
      RBC_TRY(launch_tiles_m(h->metric, src, sitems.get(), nsitems, epi, st));
    }
    // Largest candidate count, total and overflows, reduced on the device (a pageable copy
    // of all m counts would cost more than the reduction).
    RBC_TRY(ws.get(kQ + 34, 3 * 8, &p));
    unsigned long long *cs = static_cast<unsigned long long *>(p);
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, st, cs, 3);
    hipLaunchKernelGGL(cand_stats, dim3(std::min<uint32_t>(cdiv(m, 256), 64)), dim3(256), 0, st, cand_n, m, cap, cs);
    RBC_HIP_TRY(hipGetLastError());
    unsigned long long hcs[3];
    RBC_HIP_TRY(hipMemcpyAsync(hcs, cs, sizeof(hcs), hipMemcpyDeviceToHost, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
    const uint32_t mx = (uint32_t)hcs[0];
    rerank_total = (double)hcs[1];
    fallbacks += (double)hcs[2];
    cap_seen = std::max(cap_seen, mx + mx / 4);
    if (mx <= cap) break;
    if (attempt == 1) {
      set_error("candidate buffer overflow persisted (%u > %u)", mx, cap);
      return RBC_ERR_LIMIT;
    }
    cap = mx;
  }
  tm.mark();
  // 6. select
  RBC_TRY(ws.get(kQ + 19, 4 * 8, &p));
  unsigned long long *sdev = static_cast<unsigned long long *>(p);
  RBC_HIP_TRY(hipMemsetAsync(sdev, 0, 4 * 8, st));
  SelectArgs sa;
  sa.q = qd;
  sa.ldq = ldq;
  sa.xp = h->xp.get();
  sa.ldx = h->ld;
  sa.inv = h->inv.get();
  sa.n = h->n;
  sa.m = m;
  sa.d = d;
  sa.k = k;
  sa.Kp = KpUse;
  sa.probe_i = probe_i;
  // The BF16 probe's results come from probe_merge's tree-ordered re-rank (the values
  // select would compute); the FP32 probe's from the tile engine (another order).
  sa.probe_d = bprobe ? probe_d : nullptr;
  sa.probe_i1 = probe_i1;
  sa.key1 = key1;
  sa.R = R;
  sa.cert = cert;
  sa.cand_i = cand_i;
  sa.cand_n = cand_n;
  sa.cap = cap;
  sa.e_tree = tree_rel_err(tree_width(d));
  sa.out_d = od;
  sa.out_i = oi;
  sa.ldo = ldo;
  sa.stats = sdev;
  RBC_TRY(launch_select(sa, h->metric, st));
  tm.mark();
  RBC_HIP_TRY(hipStreamSynchronize(st));
  unsigned long long hs[4];
  RBC_HIP_TRY(hipMemcpy(hs, sdev, sizeof(hs), hipMemcpyDeviceToHost));
  if (hs[2]) {
    // A run of mutually unresolvable FP32 distances reached the 64th kept candidate, so
    // the order past it is not established; refuse rather than return a possibly
    // inexact answer.
    set_error("%llu queries have more than 64 candidates within the FP32 error bound of each other",
              hs[2]);
    return RBC_ERR_LIMIT;
  }
  if (stats) {
    // Pairs evaluated by the probe and the scan, from the per-list query counts.
    std::vector<uint32_t> hq(R + 1), hl(R + 1), hq1(R + 1, 0);
    void *pq0;
    RBC_TRY(ws.get(kQ + 4, (size_t(R) + 1) * 4, &pq0));
    RBC_HIP_TRY(hipMemcpy(hq.data(), pq0, (R + 1) * 4, hipMemcpyDeviceToHost));
    if (h->proposal && !bprobe) {
      void *pq1;
      RBC_TRY(ws.get(kQ + 26, (size_t(R) + 1) * 4, &pq1));
      RBC_HIP_TRY(hipMemcpy(hq1.data(), pq1, (R + 1) * 4, hipMemcpyDeviceToHost));
    }
    RBC_HIP_TRY(hipMemcpy(hl.data(), lcnt, (R + 1) * 4, hipMemcpyDeviceToHost));
    double pp = 0, ps = 0, bytes = 0;
    // Algorithmic bytes of the scan: each item reads its list once and its queries once
    // (BF16 rows of ldb, or FP32 rows of d); emitted candidates are not counted.
    const double row_bytes = sketch ? 2.0 * kSkW : bf16 ? 2.0 * h->ldb : 4.0 * d;
    const double tile = (double)scan_tile;
    for (uint32_t j = 0; j < R; ++j) {
      const double len = h->h_list_off[j + 1] - h->h_list_off[j];
      const double cq = hl[j + 1] - hl[j];
      pp += double(hq[j + 1] - hq[j]) * len + double(hq1[j + 1] - hq1[j]) * len;
      ps += cq * len;
      bytes += row_bytes * (std::ceil(cq / tile) * len + cq);
    }
    stats->pairs_probe = pp;
    stats->pairs_scan = ps;
    stats->pairs_rerank = rerank_total;
    stats->items_scan = nsitems;
    stats->lists_scanned = total_pairs;
    stats->bytes_scan = bytes;
    stats->flops_scan = sketch ? 2.0 * kSkW : bf16 ? 2.0 * h->ldb : 2.0 * d;
    stats->sketch = sketch ? 1.0 : 0.0;
    stats->guard_fp64 = (double)hs[0];
    stats->fallbacks = fallbacks;
    stats->ms_qreps = tm.ms(0, 1);
    stats->ms_probe = tm.ms(1, 2);
    stats->ms_plan = tm.ms(2, 3);
    stats->ms_scan = tm.ms(3, 4);
    stats->ms_select = tm.ms(4, 5);
    stats->ms_total = tm.ms(0, 5);
  }
  return RBC_OK;
}

int exact_query_common(rbc_exact *h, const float *qd, size_t ldq, uint32_t m, uint32_t k, int precision,
                       float *od, uint32_t *oi, uint32_t ldo, hipStream_t st, rbc_query_stats *stats) {
  RBC_TRY(exact_query_run(h, qd, ldq, m, k, precision, od, oi, ldo, st, stats));
#ifdef RBC_CHECK
  RBC_HIP_TRY(hipStreamSynchronize(st));
#endif
  return device_check();
}

}  // namespace rbc

using namespace rbc;

extern "C" int rbc_exact_set_search_params(rbc_exact *h, uint32_t probe_lists, uint32_t plane_reps) {
  RBC_ARG_CHECK(h, "NULL handle");
  RBC_ARG_CHECK(probe_lists >= 1 && probe_lists <= kProbeLists, "probe_lists=%u outside [1, %u]", probe_lists,
                kProbeLists);
  RBC_ARG_CHECK(plane_reps <= kPlaneReps, "plane_reps=%u above %u", plane_reps, kPlaneReps);
  h->probe_lists = probe_lists;
  h->plane_reps = plane_reps;
  return RBC_OK;
}

extern "C" int rbc_exact_set_proposal_mass(rbc_exact *h, float mass) {
  RBC_ARG_CHECK(h, "NULL handle");
  RBC_ARG_CHECK(mass > 0.f && mass <= 1.f, "proposal mass %g outside (0, 1]", (double)mass);
  h->prop_mass = mass;
  return RBC_OK;
}

extern "C" int rbc_exact_set_proposal(rbc_exact *h, rbc_mlp *mlp) {
  RBC_ARG_CHECK(h, "NULL handle");
  if (mlp) {
    uint32_t din, nout;
    mlp_dims(mlp, &din, &nout);
    RBC_ARG_CHECK(din == h->d && nout == h->R, "proposal network shape %u -> %u does not match d=%u, R=%u",
                  din, nout, h->d, h->R);
  }
  h->proposal = mlp;
  return RBC_OK;
}

extern "C" int rbc_exact_rep_dists(rbc_exact *h, const float *q, uint32_t m, uint32_t ldq, int precision, float *D,
                                   uint32_t ldd, float *abs_err, double *rel_err, int mem, void *stream) {
  RBC_ARG_CHECK(h, "NULL handle");
  RBC_ARG_CHECK(precision == RBC_PREC_FP32 || precision == RBC_PREC_BF16, "unknown precision %d", precision);
  RBC_ARG_CHECK(precision == RBC_PREC_FP32 || h->metric == RBC_L2, "BF16 filter needs the L2 metric");
  RBC_ARG_CHECK(D && ldd >= h->R, "NULL output or leading dimension %u < %u reps", ldd, h->R);
  const bool split = precision == RBC_PREC_BF16;
  if (rel_err) *rel_err = split ? 0.0 : fma_rel_err(h->dpad);
  if (m == 0) return RBC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevMat Q;
  RBC_TRY(Q.stage(q, m, h->d, ldq, mem, st));
  DevBuf<float> Dd, qn, qe, qn2, qerr;
  DevBuf<uint16_t> qb, qlo;
  RBC_TRY(Dd.alloc(size_t(m) * h->R));
  if (split) {
    RBC_TRY(qb.alloc(size_t(m) * h->ldb));
    RBC_TRY(qlo.alloc(size_t(m) * h->ldb));
    RBC_TRY(qn.alloc(m));
    RBC_TRY(qe.alloc(m));
    RBC_TRY(qn2.alloc(m));
    RBC_TRY(qerr.alloc(m));
  }
  RBC_TRY(rep_distances(h, Q.p, Q.ld, m, precision, Dd.get(), qb.get(), qn.get(), qe.get(), qn2.get(), qerr.get(),
                        qlo.get(), st));
  RBC_TRY(rows_out(D, size_t(ldd) * 4, Dd.get(), size_t(h->R) * 4, size_t(h->R) * 4, m, mem, st));
  if (abs_err) {
    if (split)
      RBC_TRY(rows_out(abs_err, 4, qerr.get(), 4, 4, m, mem, st));
    else if (mem == RBC_MEM_HOST)
      std::fill(abs_err, abs_err + m, 0.f);
    else
      RBC_TRY(fill_f32(abs_err, 0.f, m, st));
  }
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

extern "C" int rbc_exact_kquery(rbc_exact *h, const float *q, uint32_t m, uint32_t ldq, uint32_t k,
                                int precision, uint32_t *nn_ids, float *nn_dists, uint32_t ldo, int mem,
                                void *stream, rbc_query_stats *stats) {
  RBC_ARG_CHECK(h, "NULL handle");
  RBC_ARG_CHECK(k >= 1 && k <= RBC_MAX_K, "k=%u outside [1, %d]", k, RBC_MAX_K);
  RBC_ARG_CHECK(k <= h->n, "k=%u exceeds the database size %u", k, h->n);
  RBC_ARG_CHECK(ldo >= k, "output leading dimension %u < k", ldo);
  RBC_ARG_CHECK(precision == RBC_PREC_FP32 || precision == RBC_PREC_BF16, "unknown precision %d",
                precision);
  RBC_ARG_CHECK(precision == RBC_PREC_FP32 || h->metric == RBC_L2, "BF16 filter needs the L2 metric");
  RBC_ARG_CHECK(nn_ids && nn_dists, "NULL output");
  if (stats) *stats = rbc_query_stats{};
  if (m == 0) return RBC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevMat Q;
  RBC_TRY(Q.stage(q, m, h->d, ldq, mem, st));
  if (mem == RBC_MEM_DEVICE)
    return exact_query_common(h, Q.p, Q.ld, m, k, precision, nn_dists, nn_ids, ldo, st, stats);
  DevBuf<float> od;
  DevBuf<uint32_t> oi;
  RBC_TRY(od.alloc(size_t(m) * k));
  RBC_TRY(oi.alloc(size_t(m) * k));
  RBC_TRY(exact_query_common(h, Q.p, Q.ld, m, k, precision, od.get(), oi.get(), k, st, stats));
  RBC_TRY(rows_out(nn_dists, size_t(ldo) * 4, od.get(), size_t(k) * 4, size_t(k) * 4, m, mem, st));
  RBC_TRY(rows_out(nn_ids, size_t(ldo) * 4, oi.get(), size_t(k) * 4, size_t(k) * 4, m, mem, st));
  return RBC_OK;
}
