// BF16 MFMA distance filter of the exact query (precision RBC_PREC_BF16).
//
// The list scan computes D~ = |q~|^2 + |x~|^2 - 2 q~.x~ on BF16 copies (q~ = bf16(q),
// x~ = bf16(x)) with v_mfma_f32_16x16x32_bf16 (FP32 accumulation) and emits every point
// whose TRUE distance could still be within gamma (the probe's upper bound on the k-th
// NN distance).  With eta = |q - q~| + max_list |x - x~| (FP64, rounded up) and
// E = (gamma_K + 6u')(|q~|^2 + max_list |x~|^2) bounding the FP32 error of D~ (K the
// padded dimension, u' = 2^-23 so that no rounding mode of the accumulator is assumed),
//     |q - x| <= G  implies  D~ <= (G + eta)^2 + E,
// so the emission test is conservative; the exact FP32 re-rank decides.
//
// Work item = (one ownership list, up to 128 queries that need it).  A workgroup of four
// waves keeps the 128 queries' BF16 rows in VGPRs as MFMA A operands; each wave streams
// its own 16 points per 64-point step as B operands straight from HBM/L2 (one 16-byte
// load per lane per 32 dimensions, prefetched one step ahead).  Items are list-major and
// remapped so that each XCD takes a contiguous range: the items of a list run on one
// XCD and re-read the list from that XCD's L2.
#include "exact_query.hpp"

namespace rbc {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kBQ = 128;  // queries per item
constexpr int kBP = 64;   // points per step (16 per wave)
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

template <int KS>
__global__ __launch_bounds__(256, 2) void bf16_scan_kernel(Bf16ScanArgs a) {
  // XCD-aware remap: blocks b, b+8, b+16, ... (one XCD) take a contiguous item range.
  const uint32_t per = (a.nitems + kXCD - 1) / kXCD;
  const uint32_t item = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
  if (item >= a.nitems) return;
  const ScanItem it = a.items[item];
  __shared__ float thr[kBQ];   // t_i: emit when nx - 2 q~.x~ <= t_i
  __shared__ float sh[kBQ];    // t_i / 2 plus a margin covering FP32 rounding (pre-filter)
  __shared__ uint32_t qrow[kBQ];
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const uint32_t list = it.aux;
  if (t < kBQ) {
    uint32_t row = (uint32_t)t < it.q_count ? a.qids[it.q_begin + t] : kDummy;
    qrow[t] = row;
    float th = -INFINITY, hs = -INFINITY;
    if (row != kDummy) {
      double G = a.gam[row];
      double eta = (double)a.qe[row] + (double)a.lqe[list];
      double nq = (double)a.qn[row];
      double nxm = (double)a.lnm[list];
      double E = a.c_gemm * (nq + nxm) + 1e-30;
      double T = (G + eta) * (G + eta) + E;
      double tt = (T - nq) * (1.0 + 1e-9) + 1e-30;
      th = tt >= (double)FLT_MAX ? INFINITY : (float)tt;
      if ((double)th < tt) th = nextafterf(th, INFINITY);
      // Pre-filter: nx - 2 acc <= t  implies  acc + (t/2 + margin) >= nx/2 in FP32 when
      // the margin exceeds the rounding of both forms (|acc| <= (nq + nx)/2).
      double hh = 0.5 * tt + 0x1p-20 * (fabs(tt) + 2.0 * (nq + nxm)) + 1e-30;
      hs = hh >= (double)FLT_MAX ? INFINITY : (float)hh;
      if ((double)hs < hh) hs = nextafterf(hs, INFINITY);
    }
    thr[t] = th;
    sh[t] = hs;
  }
  __syncthreads();

  // A operands: 8 blocks of 16 queries x KS steps of 32 dims.
  bf16x8 A[8][KS];
  const uint32_t koff = (l / 16) * 8;
#pragma unroll
  for (int qb = 0; qb < 8; ++qb) {
    uint32_t row = qrow[qb * 16 + (l % 16)];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (row != kDummy)
        A[qb][s] = load8(a.qb + size_t(row) * a.ldb + s * 32 + koff);
      else
        A[qb][s] = bf16x8{};
    }
  }

  const uint32_t x0 = it.x_begin, np = it.x_count;
  const uint32_t pl = w * 16 + (l % 16);  // this lane's point within a step
  bf16x8 Bn[KS];
  {
    const bool pv = pl < np;
    const uint16_t *src = a.xb + size_t(x0 + (pv ? pl : 0)) * a.ldb + koff;
#pragma unroll
    for (int s = 0; s < KS; ++s) Bn[s] = pv ? load8(src + s * 32) : bf16x8{};
  }
  for (uint32_t p0 = 0; p0 < np; p0 += kBP) {
    bf16x8 B[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) B[s] = Bn[s];
    const uint32_t p = p0 + pl;
    const bool pv = p < np;
    // Prefetch the next step's operands.
    {
      const uint32_t pn = p + kBP;
      const bool nv = pn < np;
      const uint16_t *src = a.xb + size_t(x0 + (nv ? pn : 0)) * a.ldb + koff;
#pragma unroll
      for (int s = 0; s < KS; ++s) Bn[s] = nv ? load8(src + s * 32) : bf16x8{};
    }
    const float nx = pv ? a.xn[x0 + p] : 0.f;
    f32x4 acc[8];
#pragma unroll
    for (int qb = 0; qb < 8; ++qb) acc[qb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int qb = 0; qb < 8; ++qb)
        acc[qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[qb][s], B[s], acc[qb], 0, 0, 0);
    // Lane holds C[qb*16 + (l/16)*4 + r][point pl].  Branch-free pre-filter: one add
    // and one max per element; the exact test runs only when some lane may pass.
    float mx = -INFINITY;
#pragma unroll
    for (int qb = 0; qb < 8; ++qb) {
      const float4 s4 = *reinterpret_cast<const float4 *>(&sh[qb * 16 + (l / 16) * 4]);
      mx = fmaxf(mx, fmaxf(fmaxf(acc[qb][0] + s4.x, acc[qb][1] + s4.y),
                           fmaxf(acc[qb][2] + s4.z, acc[qb][3] + s4.w)));
    }
    const bool hit = pv && mx >= 0.5f * nx;
    if (__any(hit)) {
#pragma unroll
      for (int qb = 0; qb < 8; ++qb) {
        const int qi0 = qb * 16 + (l / 16) * 4;
        const float4 th = *reinterpret_cast<const float4 *>(&thr[qi0]);
        const float tv[4] = {th.x, th.y, th.z, th.w};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float v = __builtin_fmaf(-2.f, acc[qb][r], nx);
          if (hit && v <= tv[r]) {
            uint32_t row = qrow[qi0 + r];
            uint32_t slot = atomicAdd(&a.cand_n[row], 1u);
            if (slot < a.cap) {
              a.cand_d[size_t(row) * a.cap + slot] = v + a.qn[row];
              a.cand_i[size_t(row) * a.cap + slot] = a.perm[x0 + p];
            }
          }
        }
      }
    }
  }
}


// ---- BF16 probe -----------------------------------------------------------------------
// Same tiling as the scan; every query keeps its K' (<= 32) smallest approximate
// distances in LDS.  Candidates under the running K'-th are staged per step; one thread
// per query then folds them in by insertion.
constexpr int kPK = 32;

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
  float *out_kth;    // [m] approximate D~ of the K'-th (FLT_MAX if the list is shorter)
};

template <int KS>
__global__ __launch_bounds__(256) void bf16_probe_kernel(Bf16ProbeArgs a) {
  const ScanItem it = a.items[blockIdx.x];
  __shared__ uint32_t qrow[kBQ];
  __shared__ float qnrm[kBQ];
  __shared__ float thd[kBQ];
  __shared__ uint32_t thi[kBQ];
  __shared__ uint32_t cn[kBQ];
  __shared__ float td[kBQ][kPK];
  __shared__ uint32_t ti[kBQ][kPK];
  __shared__ float cd[kBQ][kWave];
  __shared__ uint32_t ci[kBQ][kWave];
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  if (t < kBQ) {
    uint32_t row = (uint32_t)t < it.q_count ? a.qids[it.q_begin + t] : kDummy;
    qrow[t] = row;
    qnrm[t] = row != kDummy ? a.qn[row] : 0.f;
    thd[t] = FLT_MAX;
    thi[t] = kDummy;
    cn[t] = 0;
  }
  for (int i = t; i < kBQ * kPK; i += 256) {
    td[i / kPK][i % kPK] = FLT_MAX;
    ti[i / kPK][i % kPK] = kDummy;
  }
  __syncthreads();
  bf16x8 A[8][KS];
  const uint32_t koff = (l / 16) * 8;
#pragma unroll
  for (int qb = 0; qb < 8; ++qb) {
    uint32_t row = qrow[qb * 16 + (l % 16)];
#pragma unroll
    for (int s = 0; s < KS; ++s)
      A[qb][s] = row != kDummy ? load8(a.qb + size_t(row) * a.ldb + s * 32 + koff) : bf16x8{};
  }
  const uint32_t x0 = it.x_begin, np = it.x_count;
  const uint32_t pl = w * 16 + (l % 16);
  bf16x8 Bn[KS];
  {
    const bool pv = pl < np;
    const uint16_t *src = a.xb + size_t(x0 + (pv ? pl : 0)) * a.ldb + koff;
#pragma unroll
    for (int s = 0; s < KS; ++s) Bn[s] = pv ? load8(src + s * 32) : bf16x8{};
  }
  for (uint32_t p0 = 0; p0 < np; p0 += kBP) {
    const uint32_t p = p0 + pl;
    const bool pv = p < np;
    bf16x8 B[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) B[s] = Bn[s];
    {
      const uint32_t pn = p + kBP;
      const bool nv = pn < np;
      const uint16_t *src = a.xb + size_t(x0 + (nv ? pn : 0)) * a.ldb + koff;
#pragma unroll
      for (int s = 0; s < KS; ++s) Bn[s] = nv ? load8(src + s * 32) : bf16x8{};
    }
    const float nx = pv ? a.xn[x0 + p] : 0.f;
    const uint32_t id = pv ? a.perm[x0 + p] : kDummy;
    f32x4 acc[8];
#pragma unroll
    for (int qb = 0; qb < 8; ++qb) acc[qb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int qb = 0; qb < 8; ++qb)
        acc[qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[qb][s], B[s], acc[qb], 0, 0, 0);
#pragma unroll
    for (int qb = 0; qb < 8; ++qb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int qi = qb * 16 + (l / 16) * 4 + r;
        float v = __builtin_fmaf(-2.f, acc[qb][r], nx) + qnrm[qi];
        if (pv && pair_less(v, id, thd[qi], thi[qi])) {
          uint32_t slot = atomicAdd(&cn[qi], 1u);
          cd[qi][slot] = v;
          ci[qi][slot] = id;
        }
      }
    }
    __syncthreads();
    if (t < kBQ) {
      const uint32_t n = cn[t];
      if (n) {
        float kd = thd[t];
        uint32_t ki = thi[t];
        for (uint32_t c = 0; c < n; ++c) {
          const float cv = cd[t][c];
          const uint32_t cid = ci[t][c];
          if (!pair_less(cv, cid, kd, ki)) continue;
          int pos = (int)a.Kp - 1;
          while (pos > 0 && pair_less(cv, cid, td[t][pos - 1], ti[t][pos - 1])) {
            td[t][pos] = td[t][pos - 1];
            ti[t][pos] = ti[t][pos - 1];
            --pos;
          }
          td[t][pos] = cv;
          ti[t][pos] = cid;
          kd = td[t][a.Kp - 1];
          ki = ti[t][a.Kp - 1];
        }
        thd[t] = kd;
        thi[t] = ki;
        cn[t] = 0;
      }
    }
    __syncthreads();
  }
  for (int i = t; i < kBQ * kPK; i += 256) {
    const int qi = i / kPK, j = i % kPK;
    const uint32_t row = qrow[qi];
    if (row == kDummy || j >= (int)a.Kp) continue;
    a.out_i[size_t(row) * a.Kp + j] = ti[qi][j];
    if (j == (int)a.Kp - 1) a.out_kth[row] = td[qi][j];
  }
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
    case 1: hipLaunchKernelGGL(bf16_probe_kernel<1>, dim3(nitems), dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL(bf16_probe_kernel<2>, dim3(nitems), dim3(256), 0, st, a); break;
    case 3: hipLaunchKernelGGL(bf16_probe_kernel<3>, dim3(nitems), dim3(256), 0, st, a); break;
    case 4: hipLaunchKernelGGL(bf16_probe_kernel<4>, dim3(nitems), dim3(256), 0, st, a); break;
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
