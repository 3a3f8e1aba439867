// BF16 MFMA distance filter of the exact query (precision RBC_PREC_BF16).
//
// The list scan evaluates q~.x~ on BF16 copies (q~ = bf16(q), x~ = bf16(x)) with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation) and emits every point whose TRUE
// distance could still be within G, the probe's upper bound on the k-th NN distance.
// With eta = |q - q~| + max_list |x - x~| (FP64, rounded up), |q - x| <= G implies
// |q~ - x~| <= G + eta, which the kernel tests as q~.x~ + s >= |x~|^2/2 with s folded into
// the accumulator's initial value and a margin of (gamma_K + 6u') times the absolute
// terms (u' = 2^-23, so no rounding mode of the accumulator is assumed).  The test is
// conservative; the exact FP32 re-rank decides.
//
// Scan work item = (one ownership list, up to 256 queries that need it).  The four waves
// of a workgroup keep 64 queries each as MFMA A operands in VGPRs and share the list's
// points, 32 per step, through a three-stage LDS ring filled two steps ahead.  Items are
// list-major and remapped so that each XCD takes a contiguous range: the items of a list
// run on one XCD and re-read the list from that XCD's L2.  The probe (top-K' of one
// list per query) uses 64-query items on the same LDS ring with per-lane register top-K'.
#include "exact_query.hpp"

namespace rbc {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kXCD = 8;

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// One wave per query: BF16 copy (zero padded to ldb), |q~|^2 and |q - q~| (FP64).
__global__ __launch_bounds__(256) void prep_queries(const float *q, size_t ldq, uint32_t m, uint32_t d,
                                                    uint32_t ldb, uint16_t *qb, float *qn, float *qe) {
  const uint32_t r = blockIdx.x * 4 + threadIdx.x / kWave;
  if (r >= m) return;
  const int l = lane_id();
  double nb = 0.0, e2 = 0.0;
  for (uint32_t c = l; c < ldb; c += kWave) {
    float v = c < d ? q[size_t(r) * ldq + c] : 0.f;
    uint16_t h = f32_to_bf16_rne(v);
    qb[size_t(r) * ldb + c] = h;
    double hv = (double)__uint_as_float(uint32_t(h) << 16);
    nb += hv * hv;
    double e = (double)v - hv;
    e2 += e * e;
  }
  for (int s = kWave / 2; s > 0; s >>= 1) {
    nb += __shfl_xor(nb, s, kWave);
    e2 += __shfl_xor(e2, s, kWave);
  }
  if (l == 0) {
    qn[r] = (float)nb;
    float ef = (float)sqrt(e2);
    if ((double)ef < sqrt(e2)) ef = nextafterf(ef, FLT_MAX);
    qe[r] = ef;
  }
}

struct Bf16ScanArgs {
  const ScanItem *items;
  uint32_t nitems;
  const uint32_t *qids;
  const uint16_t *qb;    // queries, BF16, ldb
  const float *qn;       // |q~|^2
  const float *qe;       // |q - q~|
  const double *gam;     // per query: upper bound on the k-th NN distance (metric units)
  const uint16_t *xb;    // permuted database, BF16, ldb
  const float *xn;       // |x~|^2 per permuted row
  const uint32_t *perm;  // permuted row -> original id
  const float *lqe;      // per list: max |x - x~|
  const float *lnm;      // per list: max |x~|^2
  uint32_t ldb;
  double c_gemm;         // gamma_K + 6u'
  float *cand_d;
  uint32_t *cand_i;
  uint32_t *cand_n;
  uint32_t cap;
};

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

// ---- BF16 list scan ------------------------------------------------------------------
// Work item = (one list, up to 256 queries).  Wave w keeps queries 64w..64w+63 as MFMA A
// operands in VGPRs; the 32 points of a step are shared by all four waves through a
// three-stage LDS ring (rows padded to 272 bytes so that the 16 lanes reading 16 different
// points at one k-offset hit distinct banks).  Each thread stages 32 bytes of a future step
// in registers: loads for step s+2 are issued while step s computes, written to LDS during
// step s+1.  One barrier per step.  Candidates go to an LDS list first (an LDS atomic does
// not wait for the global prefetch in flight) and are appended to the per-query buffers
// after the last step; entries beyond the LDS list's capacity are appended directly.
constexpr int kVQ = kBf16ScanQ;  // queries per item
constexpr int kVP = 32;          // points per step
constexpr int kRowB = 272;       // LDS bytes per point row (256 + 16 pad)
constexpr int kStages = 3;
constexpr int kCS = 1024;        // LDS candidate list entries per item

__device__ __forceinline__ int max3i(int a, int b, int c) { return max(max(a, b), c); }

template <int KS>
__global__ __launch_bounds__(256, 3) void bf16_scan_kernel(Bf16ScanArgs a) {
  const uint32_t per = (a.nitems + kXCD - 1) / kXCD;
  const uint32_t item = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
  if (item >= a.nitems) return;
  const ScanItem it = a.items[item];
  // Emission test per (query i, point x): MFMA(s_i + q~.x~) >= |x~|^2 / 2, with the
  // accumulator initialised to s_i = (T_i - |q~|^2)/2 + margin, T_i = (G_i + eta)^2.  Any x
  // with |q - x| <= G_i has |q~ - x~| <= G_i + eta, i.e. q~.x~ + (T_i - |q~|^2)/2 >= |x~|^2/2
  // in exact arithmetic; the margin covers the FP32 accumulation (gamma_{K+1} of the
  // absolute terms) and the rounding of the stored norms, so the test never drops it.
  __shared__ float sh[kVQ];
  __shared__ float qnl[kVQ];
  __shared__ uint32_t qrow[kVQ];
  __shared__ __attribute__((aligned(16))) unsigned char ring[kStages][kVP * kRowB];
  __shared__ float hring[kStages][kVP];  // 0.5 * ||bf16(x)||^2 of the step's points (INF past the end)
  __shared__ uint32_t cs_p[kCS];         // candidate: point index in the list
  __shared__ float cs_v[kCS];            // candidate: approximate squared distance
  __shared__ uint8_t cs_q[kCS];          // candidate: query slot in the item
  __shared__ uint32_t cs_n;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const uint32_t list = it.aux;
  {
    uint32_t row = (uint32_t)t < it.q_count ? a.qids[it.q_begin + t] : kDummy;
    qrow[t] = row;
    float hs = -INFINITY;
    if (row != kDummy) {
      const double G = a.gam[row];
      const double eta = (double)a.qe[row] + (double)a.lqe[list];
      const double nq = (double)a.qn[row];
      const double nxm = (double)a.lnm[list];
      const double T0 = (G + eta) * (G + eta);
      const double s0 = 0.5 * (T0 - nq);
      const double up = 0x1p-23;
      const double margin = a.c_gemm * (fabs(s0) + 0.5 * (nq + nxm)) + up * (nq + nxm) + 1e-30;
      const double hh = (s0 + margin) * (1.0 + 1e-9);
      hs = hh >= (double)FLT_MAX ? INFINITY : (float)hh;
      if ((double)hs < hh) hs = nextafterf(hs, INFINITY);
    }
    sh[t] = hs;
    qnl[t] = row != kDummy ? a.qn[row] : 0.f;
    if (t == 0) cs_n = 0;
  }
  __syncthreads();
  // This wave's queries: A operands and the initial accumulators.
  const bool wave_live = (uint32_t)(w * 64) < it.q_count;
  bf16x8 A[4][KS];
  f32x4 S[4];
  const uint32_t koff = (l / 16) * 8;
#pragma unroll
  for (int qb = 0; qb < 4; ++qb) {
    const int q0 = w * 64 + qb * 16;
    const uint32_t row = qrow[q0 + (l % 16)];
#pragma unroll
    for (int s = 0; s < KS; ++s)
      A[qb][s] = row != kDummy ? load8(a.qb + size_t(row) * a.ldb + s * 32 + koff) : bf16x8{};
    const float4 s4 = *reinterpret_cast<const float4 *>(&sh[q0 + (l / 16) * 4]);
    S[qb] = f32x4{s4.x, s4.y, s4.z, s4.w};
  }
  const uint32_t x0 = it.x_begin, np = it.x_count;
  const uint32_t nsteps = (np + kVP - 1) / kVP;
  // Staging: thread t copies bytes [32*(t%8), 32*(t%8)+32) of point t/8 of a step
  // (KS*64 bytes per row are live; the rest of the 256 is zero padding that is never read).
  const uint32_t sp = t / 8, sc = (t % 8) * 32;  // point in step, byte offset in row
  const bool stage_live = sc < (uint32_t)KS * 64;
  // Threads 0..31 also stage the squared norm of point t of the step (halved when it is
  // written to LDS, so that the load's value is not needed before the next step).
  auto gload = [&](uint32_t step, uint4 &r0, uint4 &r1, float &hn) {
    const uint32_t p = step * kVP + sp;
    if (stage_live && step < nsteps && p < np) {
      const uint4 *src = reinterpret_cast<const uint4 *>(
          reinterpret_cast<const unsigned char *>(a.xb + size_t(x0 + p) * a.ldb) + sc);
      r0 = src[0];
      r1 = src[1];
    } else {
      r0 = make_uint4(0, 0, 0, 0);
      r1 = make_uint4(0, 0, 0, 0);
    }
    const uint32_t pn = step * kVP + t;
    hn = (t < kVP && step < nsteps && pn < np) ? a.xn[x0 + pn] : INFINITY;
  };
  auto lstore = [&](int buf, const uint4 &r0, const uint4 &r1, float hn) {
    if (t < kVP) hring[buf][t] = 0.5f * hn;
    if (!stage_live) return;
    uint4 *dst = reinterpret_cast<uint4 *>(&ring[buf][sp * kRowB + sc]);
    dst[0] = r0;
    dst[1] = r1;
  };
  uint4 st0, st1;
  float sn;
  gload(0, st0, st1, sn);
  lstore(0, st0, st1, sn);
  gload(1, st0, st1, sn);
  for (uint32_t s = 0; s < nsteps; ++s) {
    __syncthreads();
    // Stage s+1 goes to LDS, the loads for s+2 start.
    lstore((s + 1) % kStages, st0, st1, sn);
    gload(s + 2, st0, st1, sn);
    if (!wave_live) continue;
    const unsigned char *buf = ring[s % kStages];
    // This lane's two points of the step: pb*16 + l%16.
    const float hx[2] = {hring[s % kStages][l % 16], hring[s % kStages][16 + l % 16]};
    f32x4 acc[4][2];
#pragma unroll
    for (int pb = 0; pb < 2; ++pb) {
      const unsigned char *rowp = buf + (pb * 16 + (l % 16)) * kRowB + koff * 2;
      bf16x8 B[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) B[ks] = *reinterpret_cast<const bf16x8 *>(rowp + ks * 64);
#pragma unroll
      for (int qb = 0; qb < 4; ++qb) acc[qb][pb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[qb][0], B[0], S[qb], 0, 0, 0);
#pragma unroll
      for (int ks = 1; ks < KS; ++ks)
#pragma unroll
        for (int qb = 0; qb < 4; ++qb)
          acc[qb][pb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[qb][ks], B[ks], acc[qb][pb], 0, 0, 0);
    }
    // Pre-filter on the float bit patterns as signed integers: for a threshold h > 0 that
    // order agrees with the float order (negative values and negative NaNs compare below,
    // positive NaNs above -- the exact test below rejects them); h == 0 (a zero-norm
    // point) always passes the pre-filter.
    bool hit = false;
#pragma unroll
    for (int pb = 0; pb < 2; ++pb) {
      int mx = max3i(__float_as_int(acc[0][pb][0]), __float_as_int(acc[0][pb][1]), __float_as_int(acc[0][pb][2]));
      mx = max3i(mx, __float_as_int(acc[0][pb][3]), __float_as_int(acc[1][pb][0]));
      mx = max3i(mx, __float_as_int(acc[1][pb][1]), __float_as_int(acc[1][pb][2]));
      mx = max3i(mx, __float_as_int(acc[1][pb][3]), __float_as_int(acc[2][pb][0]));
      mx = max3i(mx, __float_as_int(acc[2][pb][1]), __float_as_int(acc[2][pb][2]));
      mx = max3i(mx, __float_as_int(acc[2][pb][3]), __float_as_int(acc[3][pb][0]));
      mx = max3i(mx, __float_as_int(acc[3][pb][1]), __float_as_int(acc[3][pb][2]));
      mx = max(mx, __float_as_int(acc[3][pb][3]));
      const int thr = hx[pb] > 0.f ? __float_as_int(hx[pb]) : INT_MIN;
      hit = hit || mx >= thr;
    }
    if (__any(hit)) {
#pragma unroll
      for (int pb = 0; pb < 2; ++pb) {
        const uint32_t p = s * kVP + pb * 16 + (l % 16);
#pragma unroll
        for (int qb = 0; qb < 4; ++qb) {
          const int qi0 = w * 64 + qb * 16 + (l / 16) * 4;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (acc[qb][pb][r] >= hx[pb]) {
              const int qi = qi0 + r;
              const float v = qnl[qi] + 2.f * (hx[pb] - (acc[qb][pb][r] - sh[qi]));
              const uint32_t slot = atomicAdd(&cs_n, 1u);
              if (slot < (uint32_t)kCS) {
                cs_p[slot] = p;
                cs_v[slot] = v;
                cs_q[slot] = (uint8_t)qi;
              } else {
                const uint32_t row = qrow[qi];
                const uint32_t gs = atomicAdd(&a.cand_n[row], 1u);
                if (gs < a.cap) {
                  a.cand_d[size_t(row) * a.cap + gs] = v;
                  a.cand_i[size_t(row) * a.cap + gs] = a.perm[x0 + p];
                }
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  const uint32_t nc = min(cs_n, (uint32_t)kCS);
  for (uint32_t i = t; i < nc; i += 256) {
    const uint32_t row = qrow[cs_q[i]];
    const uint32_t gs = atomicAdd(&a.cand_n[row], 1u);
    if (gs < a.cap) {
      a.cand_d[size_t(row) * a.cap + gs] = cs_v[i];
      a.cand_i[size_t(row) * a.cap + gs] = a.perm[x0 + cs_p[i]];
    }
  }
}

// ---- BF16 probe -----------------------------------------------------------------------
// Work item = (one list, up to 64 queries).  The MFMA operands are swapped with respect to
// the scan: A = 16 points of the list (from the same three-stage LDS ring), B = 16 queries
// of the wave, so that lane l's accumulator holds four points of ONE query (column l%16).
// The four lanes l, l^16, l^32, l^48 see disjoint quarters of the list for their query;
// each keeps its 8 smallest approximate distances sorted in registers (strict '<': among
// equal values the earlier point stays ahead).  After the scan the four lists of a query
// are merged by one thread into the query's K' <= 32 best.
//
// The result is the exact top-K' of the list unless one quarter held more than 8 of them;
// either way every point not returned has an approximate distance >= the reported value
// kth = min(K'-th returned, 8th kept of every quarter that filled its 8 slots), which is
// the property the completeness test needs.
constexpr int kPL = 8;                 // kept per lane
constexpr int kPK = 4 * kPL;           // largest K'
constexpr int kPQ = kBf16ProbeQ;       // queries per probe item

struct Bf16ProbeArgs {
  const ScanItem *items;
  uint32_t nitems;
  const uint32_t *qids;
  const uint16_t *qb;
  const float *qn;
  const uint16_t *xb;
  const float *xn;
  const uint32_t *perm;
  uint32_t ldb, Kp;
  uint32_t *out_i;   // [m][Kp] original ids, ascending approximate distance
  float *out_kth;    // [m] lower bound of D~ over the points not returned (FLT_MAX if none)
};

template <int KS>
__global__ __launch_bounds__(256, 2) void bf16_probe_kernel(Bf16ProbeArgs a) {
  const ScanItem it = a.items[blockIdx.x];
  __shared__ uint32_t qrow[kPQ];
  __shared__ __attribute__((aligned(16))) unsigned char ring[kStages][kVP * kRowB];
  __shared__ __attribute__((aligned(16))) float xring[kStages][kVP];  // ||x~||^2 (INF past the end)
  __shared__ float lfull[kPQ][4];  // 8th kept value of each quarter that filled its list, else INF
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  if (t < kPQ) qrow[t] = (uint32_t)t < it.q_count ? a.qids[it.q_begin + t] : kDummy;
  __syncthreads();
  const uint32_t myrow = qrow[w * 16 + (l % 16)];
  const bool wave_live = (uint32_t)(w * 16) < it.q_count;
  const uint32_t koff = (l / 16) * 8;
  bf16x8 Bq[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s)
    Bq[s] = myrow != kDummy ? load8(a.qb + size_t(myrow) * a.ldb + s * 32 + koff) : bf16x8{};
  const float qn = myrow != kDummy ? a.qn[myrow] : 0.f;
  float td[kPL];
  uint32_t tp[kPL];
#pragma unroll
  for (int j = 0; j < kPL; ++j) {
    td[j] = FLT_MAX;
    tp[j] = kDummy;
  }
  const uint32_t x0 = it.x_begin, np = it.x_count;
  const uint32_t nsteps = (np + kVP - 1) / kVP;
  const uint32_t sp = t / 8, sc = (t % 8) * 32;
  const bool stage_live = sc < (uint32_t)KS * 64;
  auto gload = [&](uint32_t step, uint4 &r0, uint4 &r1, float &hn) {
    const uint32_t p = step * kVP + sp;
    if (stage_live && step < nsteps && p < np) {
      const uint4 *src = reinterpret_cast<const uint4 *>(
          reinterpret_cast<const unsigned char *>(a.xb + size_t(x0 + p) * a.ldb) + sc);
      r0 = src[0];
      r1 = src[1];
    } else {
      r0 = make_uint4(0, 0, 0, 0);
      r1 = make_uint4(0, 0, 0, 0);
    }
    const uint32_t pn = step * kVP + t;
    hn = (t < kVP && step < nsteps && pn < np) ? a.xn[x0 + pn] : INFINITY;
  };
  auto lstore = [&](int buf, const uint4 &r0, const uint4 &r1, float hn) {
    if (t < kVP) xring[buf][t] = hn;
    if (!stage_live) return;
    uint4 *dst = reinterpret_cast<uint4 *>(&ring[buf][sp * kRowB + sc]);
    dst[0] = r0;
    dst[1] = r1;
  };
  uint4 st0, st1;
  float sn;
  gload(0, st0, st1, sn);
  lstore(0, st0, st1, sn);
  gload(1, st0, st1, sn);
  for (uint32_t s = 0; s < nsteps; ++s) {
    __syncthreads();
    lstore((s + 1) % kStages, st0, st1, sn);
    gload(s + 2, st0, st1, sn);
    if (!wave_live) continue;
    const unsigned char *buf = ring[s % kStages];
#pragma unroll
    for (int pt = 0; pt < 2; ++pt) {
      const unsigned char *rowp = buf + (pt * 16 + (l % 16)) * kRowB + koff * 2;
      f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8 *>(rowp + ks * 64), Bq[ks],
                                                      acc, 0, 0, 0);
      // Lane l holds points pt*16 + (l/16)*4 + r of query l%16.
      const float4 xn4 = *reinterpret_cast<const float4 *>(&xring[s % kStages][pt * 16 + (l / 16) * 4]);
      const float xv[4] = {xn4.x, xn4.y, xn4.z, xn4.w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = __builtin_fmaf(-2.f, acc[r], xv[r]) + qn;
        if (v < td[kPL - 1]) {
          uint32_t p = s * kVP + pt * 16 + (l / 16) * 4 + r;
#pragma unroll
          for (int j = 0; j < kPL; ++j) {
            const bool lt = v < td[j];
            const float ov = td[j];
            const uint32_t op = tp[j];
            td[j] = lt ? v : ov;
            tp[j] = lt ? p : op;
            v = lt ? ov : v;
            p = lt ? op : p;
          }
        }
      }
    }
  }
  // Hand the per-lane lists to one thread per query through LDS (the ring is free now).
  __syncthreads();
  float *ld = reinterpret_cast<float *>(&ring[0][0]);              // [kPQ][4][kPL]
  uint32_t *lp = reinterpret_cast<uint32_t *>(&ring[1][0]);        // [kPQ][4][kPL]
  if (wave_live) {
    const int q = w * 16 + (l % 16), g = l / 16;
#pragma unroll
    for (int j = 0; j < kPL; ++j) {
      ld[(q * 4 + g) * kPL + j] = td[j];
      lp[(q * 4 + g) * kPL + j] = tp[j];
    }
    lfull[q][g] = tp[kPL - 1] != kDummy ? td[kPL - 1] : INFINITY;
  }
  __syncthreads();
  if (t >= kPQ || qrow[t] == kDummy) return;
  const uint32_t row = qrow[t];
  uint32_t head[4] = {0, 0, 0, 0};
  float kth = FLT_MAX;
  for (uint32_t j = 0; j < a.Kp; ++j) {
    int best = -1;
    float bv = FLT_MAX;
    uint32_t bp = kDummy;
    for (int g = 0; g < 4; ++g) {
      if (head[g] >= (uint32_t)kPL) continue;
      const float v = ld[(t * 4 + g) * kPL + head[g]];
      const uint32_t p = lp[(t * 4 + g) * kPL + head[g]];
      if (p != kDummy && (best < 0 || pair_less(v, p, bv, bp))) {
        best = g;
        bv = v;
        bp = p;
      }
    }
    if (best >= 0) ++head[best];
    a.out_i[size_t(row) * a.Kp + j] = bp != kDummy ? a.perm[x0 + bp] : kDummy;
    if (j + 1 == a.Kp) kth = bp != kDummy ? bv : FLT_MAX;
  }
  for (int g = 0; g < 4; ++g) kth = fminf(kth, lfull[t][g]);
  a.out_kth[row] = kth;
}

template <int KS>
void launch_probe(const Bf16ProbeArgs &a, uint32_t nitems, hipStream_t st) {
  hipLaunchKernelGGL((bf16_probe_kernel<KS>), dim3(nitems), dim3(256), 0, st, a);
}

}  // namespace

int bf16_prepare_queries(rbc_exact *h, const float *q, size_t ldq, uint32_t m, uint16_t *qb, float *qn,
                         float *qe, hipStream_t st) {
  hipLaunchKernelGGL(prep_queries, dim3(cdiv(m, 4)), dim3(256), 0, st, q, ldq, m, h->d, h->ldb, qb, qn, qe);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int bf16_scan(rbc_exact *h, const uint16_t *qb, const float *qn, const float *qe, const double *gam,
              const uint32_t *qlist, const ScanItem *items, uint32_t nitems, float *cand_d,
              uint32_t *cand_i, uint32_t *cand_n, uint32_t cap, hipStream_t st) {
  if (nitems == 0) return RBC_OK;
  const uint32_t KS = h->ldb / 32;
  Bf16ScanArgs a;
  a.items = items;
  a.nitems = nitems;
  a.qids = qlist;
  a.qb = qb;
  a.qn = qn;
  a.qe = qe;
  a.gam = gam;
  a.xb = h->xb.get();
  a.xn = h->xb_norm.get();
  a.perm = h->perm.get();
  a.lqe = h->list_qerr.get();
  a.lnm = h->list_nmax.get();
  a.ldb = h->ldb;
  a.c_gemm = bf16_gemm_err(h);
  a.cand_d = cand_d;
  a.cand_i = cand_i;
  a.cand_n = cand_n;
  a.cap = cap;
  const uint32_t grid = cdiv(nitems, kXCD) * kXCD;
  switch (KS) {
    case 1: hipLaunchKernelGGL(bf16_scan_kernel<1>, dim3(grid), dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL(bf16_scan_kernel<2>, dim3(grid), dim3(256), 0, st, a); break;
    case 3: hipLaunchKernelGGL(bf16_scan_kernel<3>, dim3(grid), dim3(256), 0, st, a); break;
    case 4: hipLaunchKernelGGL(bf16_scan_kernel<4>, dim3(grid), dim3(256), 0, st, a); break;
    default:
      set_error("BF16 filter supports d <= 128 (d=%u)", h->d);
      return RBC_ERR_LIMIT;
  }
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

}  // namespace rbc

namespace rbc {

int bf16_probe(rbc_exact *h, const uint16_t *qb, const float *qn, const uint32_t *qorder,
               const ScanItem *items, uint32_t nitems, uint32_t Kp, uint32_t *out_i, float *out_kth,
               hipStream_t st) {
  if (nitems == 0) return RBC_OK;
  if (Kp > (uint32_t)kPK) {
    set_error("BF16 probe keeps at most %d candidates (K'=%u)", kPK, Kp);
    return RBC_ERR_LIMIT;
  }
  Bf16ProbeArgs a;
  a.items = items;
  a.nitems = nitems;
  a.qids = qorder;
  a.qb = qb;
  a.qn = qn;
  a.xb = h->xb.get();
  a.xn = h->xb_norm.get();
  a.perm = h->perm.get();
  a.ldb = h->ldb;
  a.Kp = Kp;
  a.out_i = out_i;
  a.out_kth = out_kth;
  switch (h->ldb / 32) {
    case 1: launch_probe<1>(a, nitems, st); break;
    case 2: launch_probe<2>(a, nitems, st); break;
    case 3: launch_probe<3>(a, nitems, st); break;
    case 4: launch_probe<4>(a, nitems, st); break;
    default:
      set_error("BF16 filter supports d <= 128 (d=%u)", h->d);
      return RBC_ERR_LIMIT;
  }
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

double bf16_gemm_err(const rbc_exact *h) {
  const double up = 0x1p-23, K = (double)h->ldb;
  return K * up / (1.0 - K * up) + 6.0 * up;
}

}  // namespace rbc
