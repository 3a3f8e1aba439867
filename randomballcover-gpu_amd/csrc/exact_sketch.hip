// Projected lower-bound filter ("sketch") of the BF16 exact query (L2).
//
// For a matrix P whose rows are orthonormal, |P (q - x)| <= |q - x|: distances between
// projections bound the true distances from below.  The build takes P = the top
// kSkDims principal directions of a sample of the database (FP64 covariance on the GPU,
// Jacobi eigenvectors on the host, re-orthonormalised, stored in FP32), and keeps for
// every point the 32-coordinate BF16 row
//     x^ = (y~_0 .. y~_27, -n_hi, -n_lo, 1, 1),   y = P (x - mu),  y~ = bf16(y),
// with n_hi + n_lo <= |y~|^2 / 2 (both rounded down).  A query gets (y~_q, 1, 1, s_hi,
// s_lo) with s_hi + s_lo >= s (rounded up), s the emission threshold of its (query, list)
// pair, so that ONE v_mfma_f32_16x16x32_bf16 product x^ . q^ equals
//     y~_q . y~_x - |y~_x|^2 / 2 + s,
// and a point is emitted when that is >= 0: the threshold test needs no vector-ALU work
// besides a max over the accumulators.  The scan thus costs 64 FLOPs per pair instead of
// 2 ldb; the exact FP32 re-rank decides every result, as with the full-dimension filter.
//
// Exactness.  With s_P >= |P|_2 (from P P^T - I, FP64), any x with |q - x| <= G has
// |P (q - x)| <= s_P G.  The computed projection y^ (FP32, one FMA per coordinate after an
// FP32 subtraction of mu) satisfies |y^_j - (P (x - mu))_j| <= gamma_{d+1} sum_i |P_ji|
// |x_i - mu_i| (evaluated in FP64 per coordinate), and e_x = that vector bound + |y^ - y~|
// (FP64, rounded up), so |y~_q - y~_x| <= s_P G + e_q + e_x =: G + eta in the notation of
// the full-dimension filter, and the same threshold algebra applies with the stored norms
// |y~|^2 and the per-list maxima of e_x and |y~_x|^2.
//
// The filter pays off only where the data lie close to a low-dimensional subspace: the
// bound loses about 2 x (variance outside the kSkDims directions) in squared distance.
// The build decides (RBC_SKETCH_AUTO) by comparing that loss with the mean squared
// distance of a point to its representative.
#include <algorithm>
#include <cmath>
#include <vector>

#include "exact_query.hpp"
#include "exact_sketch.hpp"
#include "scan_sched.hpp"

namespace rbc {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4;
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(3))) const i32x4 lds_i32x4;

constexpr uint32_t kSkSample = 65536;  // rows sampled for the mean and the covariance
constexpr uint32_t kSkChunk = 2048;    // sample rows per covariance partial


// ---- build: mean, covariance and the representative distances of a row sample ---------
// Sample row i (i < S) is permuted row floor(i n / S).
__device__ __forceinline__ uint32_t sample_row(uint32_t i, uint32_t S, uint32_t n) {
  return (uint32_t)(((unsigned long long)i * n) / S);
}

// Block (c0/64, ch): coordinates c0 .. c0+63 over the chunk's sample rows; thread
// (part = t / 64, c = t % 64) sums rows part, part + 4, ... of the chunk in FP64, and the
// four parts are added in a fixed order into part_out[ch][col].
__global__ __launch_bounds__(256) void sk_mean_part(const float *xp, size_t ld, uint32_t n, uint32_t S, uint32_t d,
                                                    double *part_out) {
  __shared__ double part[4][64];
  const int t = threadIdx.x, c = t % 64, pt = t / 64;
  const uint32_t col = blockIdx.x * 64 + c, ch = blockIdx.y;
  const uint32_t s0 = ch * kSkChunk, s1 = min(S, s0 + kSkChunk);
  double s = 0.0;
  if (col < d)
    for (uint32_t i = s0 + pt; i < s1; i += 4) s += (double)xp[size_t(sample_row(i, S, n)) * ld + col];
  part[pt][c] = s;
  __syncthreads();
  if (pt == 0 && col < d) part_out[size_t(ch) * d + col] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// mu[col] = sum over chunks (in chunk order) / S.
__global__ void sk_mean_sum(const double *part, uint32_t nch, uint32_t d, uint32_t S, double *mu) {
  const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= d) return;
  double s = 0.0;
  for (uint32_t c = 0; c < nch; ++c) s += part[size_t(c) * d + col];
  mu[col] = s / (double)S;
}

// Partial covariance sums: block (tile ti, tile tj, chunk ch) adds, over the chunk's
// sample rows, (x_i - mu_i)(x_j - mu_j) for the 32 x 32 entries of the tile into
// part[ch][i][j]; thread t holds row ti*32 + t/8 and columns tj*32 + (t%8)*4 .. +3.
__global__ __launch_bounds__(256) void sk_cov_part(const float *xp, size_t ld, uint32_t n, uint32_t S, uint32_t d,
                                                   const double *mu, double *part) {
  __shared__ double si[64][33], sj[64][33];
  const int t = threadIdx.x;
  const uint32_t i0 = blockIdx.x * 32, j0 = blockIdx.y * 32, ch = blockIdx.z;
  const uint32_t r = t / 8, c4 = (t % 8) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const uint32_t s0 = ch * kSkChunk, s1 = min(S, s0 + kSkChunk);
  for (uint32_t b = s0; b < s1; b += 64) {
    __syncthreads();
    for (int e = t; e < 64 * 32; e += 256) {
      const uint32_t sr = e / 32, cc = e % 32, row = b + sr;
      double vi = 0.0, vj = 0.0;
      if (row < s1) {
        const float *xr = xp + size_t(sample_row(row, S, n)) * ld;
        if (i0 + cc < d) vi = (double)xr[i0 + cc] - mu[i0 + cc];
        if (j0 + cc < d) vj = (double)xr[j0 + cc] - mu[j0 + cc];
      }
      si[sr][cc] = vi;
      sj[sr][cc] = vj;
    }
    __syncthreads();
    for (int sr = 0; sr < 64; ++sr) {
      const double a = si[sr][r];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] += a * sj[sr][c4 + u];
    }
  }
  const uint32_t i = i0 + r;
  if (i < d)
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (j0 + c4 + u < d) part[(size_t(ch) * d + i) * d + j0 + c4 + u] = acc[u];
}

// C[i][j] = sum over chunks (in chunk order) / S.
__global__ void sk_cov_sum(const double *part, uint32_t nch, uint32_t d, uint32_t S, double *C) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d * d) return;
  double s = 0.0;
  for (uint32_t c = 0; c < nch; ++c) s += part[size_t(c) * d * d + e];
  C[e] = s / (double)S;
}

// Squared FP64 distance of each sample row to its owning representative (the scale the
// lower bound's loss is compared with).
__global__ void sk_owner_d2(const float *xp, size_t ld, uint32_t n, uint32_t S, uint32_t d, const uint32_t *perm,
                            const uint32_t *owner, const float *reps, double *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  const uint32_t p = sample_row(i, S, n);
  const float *xr = xp + size_t(p) * ld, *rr = reps + size_t(owner[perm[p]]) * ld;
  double s = 0.0;
  for (uint32_t c = 0; c < d; ++c) {
    const double t = (double)xr[c] - (double)rr[c];
    s += t * t;
  }
  out[i] = s;
}

// ---- projection of rows (database at build, queries per call) --------------------------
// 32 rows per block, in 4 rounds of 8 rows with 32 threads per row (P is staged once per
// block); thread j < kSkDims computes y_j = sum_i P[j][i] (x_i - mu_i) with one FMA per
// coordinate in coordinate order (FP32, after an FP32 subtraction), and the bound
// gamma_{d+1} a_j, a_j = sum_i |P_ji| |x_i - mu_i|: a_j is summed in FP32 from |x_i - mu_i|
// rounded up, and a sum of d non-negative terms with one FMA each is within gamma_d of its
// exact value, so (1 + 2.02 d 2^-24) times it bounds a_j from above.  Outputs: the BF16 row
// (database: -n_hi, -n_lo, 1, 1 in the last four coordinates; queries: 1, 1, 0, 0), e =
// |y^ - y~| + |bound vector| rounded up, and |y~|^2 (database: rounded up; queries: rounded
// to nearest).
constexpr uint32_t kSkProjRows = 32;

__global__ __launch_bounds__(256) void sk_project(const float *x, size_t ldx, uint32_t rows, uint32_t d,
                                                  const float *P, uint32_t ldP, const float *mu, double gam_d1,
                                                  int database, uint16_t *out, float *err, float *n2) {
  // P by rows (coordinate i of row j at PT[j][i], rows padded by 4 floats so that the 16
  // lanes of a 16-byte read pass fall on distinct banks), read 4 coordinates at a time.
  constexpr int kPad = kSkMaxD + 4;
  __shared__ __attribute__((aligned(16))) float PT[kSkW][kPad];
  __shared__ __attribute__((aligned(16))) float xs[8][kSkMaxD];
  __shared__ __attribute__((aligned(16))) float xa[8][kSkMaxD];
  const int t = threadIdx.x, r = t / kSkW, j = t % kSkW;
  for (uint32_t e = t; e < kSkW * kSkMaxD; e += 256) {
    const uint32_t jj = e / kSkMaxD, i = e % kSkMaxD;
    PT[jj][i] = jj < (uint32_t)kSkDims && i < d ? P[size_t(jj) * ldP + i] : 0.f;
  }
  const double afac = 1.0 + 2.02 * (double)d * 0x1p-24;
  for (uint32_t round = 0; round < kSkProjRows / 8; ++round) {
    const uint32_t row = blockIdx.x * kSkProjRows + round * 8 + r;
    __syncthreads();  // P staged; the previous round's rows read
    for (uint32_t i = j; i < kSkMaxD; i += kSkW) {
      const float v = row < rows && i < d ? x[size_t(row) * ldx + i] : 0.f;
      xs[r][i] = i < d ? v - mu[i] : 0.f;
      // |x_i - mu_i| rounded up, for the bound (the FP32 difference may round down).
      xa[r][i] = i < d ? f32_up(fabs((double)v - (double)mu[i])) : 0.f;
    }
    __syncthreads();
    float y = 0.f, af = 0.f;
    for (uint32_t i = 0; i < d; i += 4) {
      const float4 p4 = *reinterpret_cast<const float4 *>(&PT[j][i]);
      const float4 x4 = *reinterpret_cast<const float4 *>(&xs[r][i]);
      const float4 a4 = *reinterpret_cast<const float4 *>(&xa[r][i]);
      y = __builtin_fmaf(p4.x, x4.x, y);
      y = __builtin_fmaf(p4.y, x4.y, y);
      y = __builtin_fmaf(p4.z, x4.z, y);
      y = __builtin_fmaf(p4.w, x4.w, y);
      af = __builtin_fmaf(fabsf(p4.x), a4.x, af);
      af = __builtin_fmaf(fabsf(p4.y), a4.y, af);
      af = __builtin_fmaf(fabsf(p4.z), a4.z, af);
      af = __builtin_fmaf(fabsf(p4.w), a4.w, af);
    }
    double q2 = 0.0, b2 = 0.0, y2 = 0.0;
    uint16_t h = 0;
    if (j < kSkDims) {
      h = bf16_rne(y);
      const double hv = (double)bf16_val(h), dq = (double)y - hv, bj = gam_d1 * ((double)af * afac);
      q2 = dq * dq;
      b2 = bj * bj;
      y2 = hv * hv;
    }
    // Sums over the row's 32 threads (lanes 32 (r % 2) .. +31 of the wave).
    for (int o = kSkW / 2; o > 0; o >>= 1) {
      q2 += __shfl_xor(q2, o, kSkW);
      b2 += __shfl_xor(b2, o, kSkW);
      y2 += __shfl_xor(y2, o, kSkW);
    }
    if (row >= rows) continue;
    if (j >= kSkDims) {
      if (database) {
        // -n_hi, -n_lo with n_hi + n_lo <= |y~|^2 / 2.
        const double N = 0.5 * y2;
        const uint16_t nh = bf16_down_pos(f32_down(N));
        const double rem = N - (double)bf16_val(nh);
        const uint16_t nl = rem > 0.0 ? bf16_down_pos(f32_down(rem)) : (uint16_t)0;
        h = j == kSkDims ? (uint16_t)(nh | 0x8000u) : j == kSkDims + 1 ? (uint16_t)(nl | 0x8000u) : kBf16One;
      } else {
        h = j < kSkDims + 2 ? kBf16One : (uint16_t)0;
      }
    }
    out[size_t(row) * kSkW + j] = h;
    if (j == 0) {
      err[row] = f32_up((sqrt(q2) + sqrt(b2)) * (1.0 + 1e-12));
      n2[row] = database ? f32_up(y2) : (float)y2;
    }
  }
}

// One block per list: the largest e and the largest |y~|^2 of its rows.
__global__ __launch_bounds__(256) void sk_list_stats(const uint32_t *off, const float *err, const float *n2,
                                                     float *lqe, float *lnm) {
  __shared__ float se[256], sn[256];
  const uint32_t j = blockIdx.x;
  float me = 0.f, mn = 0.f;
  for (uint32_t p = off[j] + threadIdx.x; p < off[j + 1]; p += 256) {
    me = fmaxf(me, err[p]);
    mn = fmaxf(mn, n2[p]);
  }
  se[threadIdx.x] = me;
  sn[threadIdx.x] = mn;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] = fmaxf(se[threadIdx.x], se[threadIdx.x + s]);
      sn[threadIdx.x] = fmaxf(sn[threadIdx.x], sn[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    lqe[j] = se[0];
    lnm[j] = sn[0];
  }
}

// ---- the list scan on sketches ----------------------------------------------------------
// Work item = (one ownership list, up to 64 QB queries).  Wave w keeps queries 16 QB w ..
// as QB MFMA B operands (one 16-byte fragment each: K = 32, the query's threshold in
// coordinates 30 and 31); the list's sketch rows, the A operands, stream through an
// NST-stage LDS ring, 16 PB points per step, filled NST - 1 steps ahead by direct
// global->LDS loads (PB / 4 loads of 1 KB per wave and step).  A point block (16 points)
// is one ds_read_b128 per lane feeding QB MFMAs whose accumulators start at 0 (the next
// block's fragment is read while they run); lane l then holds points 4 (l/16) .. +3 of
// query l % 16 of every block.  The emission test is acc >= 0 on the float bits as signed
// integers (a max3 tree over the block's 4 QB accumulators); a point block with a hit
// re-tests its values one by one and appends (point, query slot) to an LDS list that is
// flushed to the per-query candidate buffers after the item's last step.
// Ring rows are 64 bytes (4 chunks of 16); chunk c of row r sits in slot c ^ swz(r) with
// swz(r) = ((r >> 2) ^ (r >> 1)) & 3: of the XOR patterns timed on a ds_read_b128 of 16
// rows x 4 chunks it gave the fewest bank conflicts (1.3 ns per wave read per CU against
// 1.8 ns for ((r >> 1) & 3) ^ ((r >> 3) & 1)).
// Measured on low-rank data (n = 4M, d = 128, 16-dimensional, 131,072 queries, 2,048
// lists): 8 blocks of 16 queries per wave (4 waves per SIMD) beat 16
// (2 waves per SIMD); steps of 128 points beat 64 (half the barriers); waves loading
// their fragments from L2 without a shared ring were 15-20 % slower.

__device__ __forceinline__ uint32_t sk_swz(uint32_t r) { return ((r >> 2) ^ (r >> 1)) & 3u; }

__device__ __forceinline__ uint32_t lds_addr(const void *p) { return (uint32_t)(uintptr_t)(const lds_void *)p; }

// One direct global->LDS load of 16 bytes per lane (global_load_lds_dwordx4: lane l
// writes LDS bytes [16 l, 16 l + 16) after the wave-uniform address in M0), issued from
// asm so that the compiler neither drains it before unrelated instructions nor moves LDS
// reads across it; M0 is saved and restored inside the block (s_nop: the wait state
// between writing M0 and the LDS DMA that reads it).
__device__ __forceinline__ void sk_dma(const void *base, uint32_t off, uint32_t m0) {
  uint32_t keep;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %2\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %1, %3\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(keep)
      : "v"(off), "s"(m0), "s"(base)
      : "memory");
}

// vmcnt(n) with the other counters left alone (gfx9 encoding).
template <int N>
__device__ __forceinline__ void sk_wait_vm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

__device__ __forceinline__ int imax3(int a, int b, int c) { return max(max(a, b), c); }


// Waits until this wave has at most n ring loads outstanding (n < 8).
__device__ __forceinline__ void sk_wait_upto(uint32_t n) {
  switch (n) {
    case 0: sk_wait_vm<0>(); break;
    case 1: sk_wait_vm<1>(); break;
    case 2: sk_wait_vm<2>(); break;
    case 3: sk_wait_vm<3>(); break;
    case 4: sk_wait_vm<4>(); break;
    case 5: sk_wait_vm<5>(); break;
    default: sk_wait_vm<6>(); break;
  }
}

// QB: 16-query blocks per wave; CS: LDS candidate entries; PF: read the next point
// block's fragment ahead; PB: point blocks (of 16) per step (4 or 8: each wave loads PB/4
// KB per step); NST: ring stages (prefetch distance NST - 1 steps).
template <int QB, int CS, bool PF, int PB, int NST, int NW>
__global__ __launch_bounds__(64 * NW) void sk_scan_kernel(SkScanArgs a) {
  constexpr int QW = 16 * QB, VQ = NW * QW, NT = 64 * NW;
  constexpr int kSkPB = PB, kSkVP = 16 * PB, kSkStages = NST, kSkStageB = kSkVP * 64, LW = PB / NW;
  static_assert(PB % NW == 0 && LW * (NST - 2) <= 6, "ring shape");
  __shared__ uint32_t qrow[VQ];
  __shared__ uint32_t sthr[VQ];
  __shared__ __attribute__((aligned(16))) unsigned char ring[kSkStages * kSkStageB];
  __shared__ uint32_t cs_p[CS];
  __shared__ uint16_t cs_q[CS];
  __shared__ uint32_t cs_n;
  __shared__ uint32_t s_item;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const uint32_t xcd = blockIdx.x % kXCD;
  const uint32_t lo = a.bounds[xcd], hi = a.bounds[xcd + 1];
  const uint32_t wu = __builtin_amdgcn_readfirstlane((uint32_t)w);
  // Ring loads: wave w fills block w (rows 16 w .. +15) of a stage; lane l writes row l/4,
  // slot l%4, with the row's chunk that belongs in that slot.
  const uint32_t m0w = lds_addr(ring) + wu * 1024;
  const uint32_t drow = (uint32_t)l / 4, dchunk = ((uint32_t)l % 4) ^ sk_swz(drow);
  // Ring reads: lane l reads row l%16, chunk l/16 of a point block.
  const uint32_t roff = ((uint32_t)l % 16) * 64 + ((((uint32_t)l / 16) ^ sk_swz((uint32_t)l % 16)) * 16);
  // An item index from this XCD's range, else (when it is exhausted) from the others, in
  // XCD order after this one (so the XCDs finish together); kDummy when all are done.
  auto take = [&](uint32_t v) -> uint32_t {
    if (v < hi) return v;
    for (uint32_t dd = 1; dd < (uint32_t)kXCD; ++dd) {
      const uint32_t y = (xcd + dd) % kXCD;
      const uint32_t u = a.bounds[y] + atomicAdd(&a.next[y], 1u);
      if (u < a.bounds[y + 1]) return u;
    }
    return kDummy;
  };
  if (t == 0) s_item = take(lo + atomicAdd(&a.next[xcd], 1u));
  __syncthreads();
  for (;;) {
    const uint32_t item = s_item;
    if (item == kDummy) break;
    // The following item is requested now and read after this one (its atomic's latency
    // hides behind the item's work).
    uint32_t nxt = 0;
    if (t == 0) nxt = lo + atomicAdd(&a.next[xcd], 1u);
    const ScanItem it = a.items[item];
    const uint32_t x0 = __builtin_amdgcn_readfirstlane(it.x_begin);
    const uint32_t np = __builtin_amdgcn_readfirstlane(it.x_count);
    const uint32_t nsteps = (np + kSkVP - 1) / kSkVP;
    // Rows past the list end load clamped copies of its last row (every lane issues its
    // load, so each wave issues exactly one load per step); the exact test drops them.
    auto issue = [&](uint32_t step) {
      const uint32_t p0 = step * kSkVP, last = np - 1 - p0;
      const uint16_t *src = a.xs + size_t(x0 + p0) * kSkW;
#pragma unroll
      for (int j = 0; j < LW; ++j)
        sk_dma(src, min((wu + NW * j) * 16 + drow, last) * 64 + dchunk * 16,
               m0w + j * (NW * 1024) + (step % kSkStages) * kSkStageB);
    };
#pragma unroll
    for (int s0 = 0; s0 < NST - 1; ++s0)
      if ((uint32_t)s0 < nsteps) issue(s0);
    // Query slots of wave w: QW w .. QW w + QW - 1; lane l fetches the ids of slots
    // QW w + 64 h + l and computes their thresholds for this list.
    uint32_t myq[QB / 4];
#pragma unroll
    for (int h = 0; h < QB / 4; ++h) {
      const uint32_t slot = QW * w + 64 * h + l;
      const bool live = slot < it.q_count;
      myq[h] = live ? a.qids[it.q_begin + slot] : kDummy;
      qrow[slot] = myq[h];
      sthr[slot] = live ? a.qthr[myq[h]] : (uint32_t)kBf16NegInf;
    }
    const bool wave_live = (uint32_t)(w * QW) < it.q_count;
    bf16x8 B[QB];
    const uint32_t koff = ((uint32_t)l / 16) * 8;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      const uint32_t row = (uint32_t)__shfl((int)myq[qb / 4], (qb % 4) * 16 + (l % 16), kWave);
      i32x4 v = i32x4{0, 0, 0, 0};
      if (row != kDummy) v = *reinterpret_cast<const i32x4 *>(a.qs + size_t(row) * kSkW + koff);
      B[qb] = __builtin_bit_cast(bf16x8, v);
    }
    if (t == 0) cs_n = 0;
    __syncthreads();
    // Lanes 48..63 hold coordinates 24..31: coordinates 30, 31 carry the threshold of
    // the (query, list) pair (dummy slots: -inf, so that they never emit).
    if (l >= 48) {
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        i32x4 v = __builtin_bit_cast(i32x4, B[qb]);
        v[3] = (int)sthr[QW * w + 16 * qb + (l % 16)];
        B[qb] = __builtin_bit_cast(bf16x8, v);
      }
    }
    // The query operands must have arrived before the first MFMA; waiting here (for the
    // ring's first loads too) keeps the compiler from placing that wait inside the loop.
    sk_wait_vm<0>();
    for (uint32_t s = 0; s < nsteps; ++s) {
      // This wave's loads of step s are done when at most those of the following NST - 2
      // steps (those that exist) are outstanding; the barrier then publishes every wave's.
      sk_wait_upto(LW * min((uint32_t)(NST - 2), nsteps - 1 - s));
      asm volatile("s_barrier" ::: "memory");
      if (s + NST - 1 < nsteps) issue(s + NST - 1);
      if (!wave_live) continue;
      const uint32_t so = (s % kSkStages) * kSkStageB;
      const uint32_t left = np - s * kSkVP;
      const uint32_t rbase = lds_addr(ring) + so + roff;
      i32x4 rawn = *reinterpret_cast<const lds_i32x4 *>((uintptr_t)rbase);
#pragma unroll
      for (int pb = 0; pb < kSkPB; ++pb) {
        if ((uint32_t)pb * 16 >= left) break;
        i32x4 raw;
        if constexpr (PF) {
          // The next block's fragment is read while this block's MFMAs run.
          raw = rawn;
          if (pb + 1 < kSkPB) rawn = *reinterpret_cast<const lds_i32x4 *>((uintptr_t)(rbase + (pb + 1) * 1024));
        } else {
          raw = pb == 0 ? rawn : *reinterpret_cast<const lds_i32x4 *>((uintptr_t)(rbase + pb * 1024));
        }
        const bf16x8 A = __builtin_bit_cast(bf16x8, raw);
        f32x4 acc[QB];
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[qb], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        __builtin_amdgcn_s_setprio(0);
        // max over the 4 QB accumulators as signed integers: >= 0 iff some float >= +0.
        int mx[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) mx[c] = INT_MIN;
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) {
          mx[qb % 4] = imax3(mx[qb % 4], __float_as_int(acc[qb][0]), __float_as_int(acc[qb][1]));
          mx[qb % 4] = imax3(mx[qb % 4], __float_as_int(acc[qb][2]), __float_as_int(acc[qb][3]));
        }
        const int m = max(imax3(mx[0], mx[1], mx[2]), mx[3]);
        if (__any(m >= 0)) {
          const uint32_t p0 = s * kSkVP + pb * 16 + ((uint32_t)l / 16) * 4;
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) {
            const int mq = max(imax3(__float_as_int(acc[qb][0]), __float_as_int(acc[qb][1]),
                                     __float_as_int(acc[qb][2])),
                               __float_as_int(acc[qb][3]));
            if (mq < 0) continue;
            const uint32_t qi = QW * w + 16 * qb + ((uint32_t)l % 16);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const uint32_t p = p0 + r;
              if (__float_as_int(acc[qb][r]) >= 0 && p < np) {
                const uint32_t slot = atomicAdd(&cs_n, 1u);
                if (slot < (uint32_t)CS) {
                  cs_p[slot] = p;
                  cs_q[slot] = (uint16_t)qi;
                } else {
                  const uint32_t qr = qrow[qi];
                  const uint32_t gs = atomicAdd(&a.cand_n[qr], 1u);
                  if (gs < a.cap) a.cand_i[size_t(qr) * a.cap + gs] = a.perm[RBC_CHECKED(x0 + p, a.n)];
                }
              }
            }
          }
        }
      }
    }
    __syncthreads();
    const uint32_t nc = min(cs_n, (uint32_t)CS);
    for (uint32_t i = t; i < nc; i += NT) {
      const uint32_t qr = qrow[cs_q[i]];
      const uint32_t gs = atomicAdd(&a.cand_n[qr], 1u);
      if (gs < a.cap) a.cand_i[size_t(qr) * a.cap + gs] = a.perm[RBC_CHECKED(x0 + cs_p[i], a.n)];
    }
    if (t == 0) s_item = take(nxt);
    // The next item rewrites qrow, sthr and cs_*, after every thread has read them.
    __syncthreads();
  }
}

// ---- the probe on sketches -------------------------------------------------------------
// Work item = (one list, up to 64 probe entries; entry e belongs to query e / P).  Wave w
// keeps entries 16 w .. 16 w + 15 as one MFMA B operand: the query's sketch row with
// coordinates 30, 31 set to about -|y~_q|^2 / 2 (the same value for all of a query's
// entries), so that the MFMA gives acc = -D~/2, D~ the squared sketch distance (up to a
// per-query constant).  acc <= 0 for every point but a coincident one, and the float bits
// of non-positive values order by magnitude as unsigned integers, so a point's key is
// its accumulator's bits with the low bits replaced by the point's slot (point block,
// row) -- one v_and_or per pair -- and each lane keeps its PL smallest keys with the
// branch-free median insertion.  The list's sketch rows come through the scan's ring
// (direct global->LDS loads, 128 points per step).  After the last step one thread per
// entry merges the four lanes that saw disjoint quarters of the list into the entry's
// ascending list of up to 4 PL (id, D~) pairs.  Only a bound is taken from them: the
// caller re-ranks its best candidates exactly.
constexpr uint32_t kSkNoKey = 0xFF800000u;   // keys at or above this hold no point

__device__ __forceinline__ uint32_t sk_med3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// Inserts key into the ascending list t (the largest drops out): slot j takes the median
// of (t[j-1], key, t[j]).
template <int PL>
__device__ __forceinline__ void sk_insert(uint32_t (&t)[PL], uint32_t key) {
#pragma unroll
  for (int j = PL - 1; j > 0; --j) t[j] = sk_med3(t[j - 1], key, t[j]);
  t[0] = min(t[0], key);
}

// PL: keys kept per lane (entries hold 4 PL candidates).
template <int PB, int NST, int PL>
__global__ __launch_bounds__(256) void sk_probe_kernel(SkProbeArgs a) {
  constexpr int VP = 16 * PB, StageB = VP * 64, LW = PB / 4;
  static_assert(PB % 4 == 0 && LW * (NST - 2) <= 6, "ring shape");
  constexpr int kE = 64;  // entries per item
  __shared__ uint32_t qent[kE];
  __shared__ __attribute__((aligned(16))) unsigned char ring[NST * StageB];
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  // Workgroup b runs on XCD b % kXCD; each XCD takes a contiguous run of the list-major
  // items, so the items of one list re-read it from the same L2.
  const uint32_t per_xcd = (a.nitems + kXCD - 1) / kXCD;
  const uint32_t item = (blockIdx.x % kXCD) * per_xcd + blockIdx.x / kXCD;
  if (item >= a.nitems) return;
  const ScanItem it = a.items[item];
  const uint32_t x0 = __builtin_amdgcn_readfirstlane(it.x_begin);
  const uint32_t np = __builtin_amdgcn_readfirstlane(it.x_count);
  const uint32_t nsteps = (np + VP - 1) / VP;
  const uint32_t wu = __builtin_amdgcn_readfirstlane((uint32_t)w);
  const uint32_t m0w = lds_addr(ring) + wu * 1024;
  const uint32_t drow = (uint32_t)l / 4, dchunk = ((uint32_t)l % 4) ^ sk_swz(drow);
  const uint32_t roff = ((uint32_t)l % 16) * 64 + ((((uint32_t)l / 16) ^ sk_swz((uint32_t)l % 16)) * 16);
  auto issue = [&](uint32_t step) {
    const uint32_t p0 = step * VP, last = np - 1 - p0;
    const uint16_t *src = a.xs + size_t(x0 + p0) * kSkW;
#pragma unroll
    for (int j = 0; j < LW; ++j)
      sk_dma(src, min((wu + 4 * j) * 16 + drow, last) * 64 + dchunk * 16, m0w + j * 4096 + (step % NST) * StageB);
  };
#pragma unroll
  for (int s0 = 0; s0 < NST - 1; ++s0)
    if ((uint32_t)s0 < nsteps) issue(s0);
  if (t < kE) qent[t] = (uint32_t)t < it.q_count ? a.qids[it.q_begin + t] : kDummy;
  __syncthreads();
  const uint32_t myent = qent[w * 16 + (l % 16)];
  const uint32_t myrow = myent != kDummy ? myent / a.P : kDummy;
  const bool wave_live = (uint32_t)(w * 16) < it.q_count;
  i32x4 bv = i32x4{0, 0, 0, 0};
  if (myrow != kDummy) bv = *reinterpret_cast<const i32x4 *>(a.qs + size_t(myrow) * kSkW + ((uint32_t)l / 16) * 8);
  if (l >= 48) {
    uint32_t sp = (uint32_t)kBf16NegInf;
    if (myrow != kDummy) {
      const float s = -0.5f * a.qn[myrow];
      const uint16_t sh = bf16_rne(s);
      sp = ((uint32_t)bf16_rne(s - bf16_val(sh)) << 16) | sh;
    }
    bv[3] = (int)sp;
  }
  const bf16x8 B = __builtin_bit_cast(bf16x8, bv);
  const uint32_t vmask = ~((1u << a.idx_bits) - 1u);
  uint32_t tk[PL];
#pragma unroll
  for (int j = 0; j < PL; ++j) tk[j] = 0xFFFFFFFFu;
  sk_wait_vm<0>();
  for (uint32_t s = 0; s < nsteps; ++s) {
    sk_wait_upto(LW * min((uint32_t)(NST - 2), nsteps - 1 - s));
    asm volatile("s_barrier" ::: "memory");
    if (s + NST - 1 < nsteps) issue(s + NST - 1);
    if (!wave_live) continue;
    const uint32_t rbase = lds_addr(ring) + (s % NST) * StageB + roff;
    const uint32_t left = np - s * VP;
#pragma unroll
    for (int pb = 0; pb < PB; ++pb) {
      if ((uint32_t)pb * 16 >= left) break;
      const i32x4 raw = *reinterpret_cast<const lds_i32x4 *>((uintptr_t)(rbase + pb * 1024));
      const f32x4 acc =
          __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, raw), B, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      const uint32_t slot0 = (s * PB + pb) * 4;
      if ((uint32_t)pb * 16 + 16 <= left) {
#pragma unroll
        for (int r = 0; r < 4; ++r) sk_insert(tk, (__float_as_uint(acc[r]) & vmask) | (slot0 + r));
      } else {
        // The list's last block: rows past its end hold copies of its last row.
        const uint32_t p0 = (uint32_t)pb * 16 + ((uint32_t)l / 16) * 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t key = p0 + r < left ? (__float_as_uint(acc[r]) & vmask) | (slot0 + r) : 0xFFFFFFFFu;
          sk_insert(tk, key);
        }
      }
    }
  }
  // Hand the per-lane lists to one thread per entry through LDS (the ring is free now).
  __syncthreads();
  uint32_t *lk = reinterpret_cast<uint32_t *>(ring);  // [kE][4][PL]
  if (wave_live) {
    const int e = w * 16 + (l % 16), g = l / 16;
#pragma unroll
    for (int j = 0; j < PL; ++j) lk[(e * 4 + g) * PL + j] = tk[j];
  }
  __syncthreads();
  if (t >= kE || qent[t] == kDummy) return;
  const uint32_t ent = qent[t];
  const uint32_t smask = ~vmask;
  uint32_t head[4] = {0, 0, 0, 0};
  for (uint32_t j = 0; j < a.Kp; ++j) {
    int best = -1;
    uint32_t bk = 0xFFFFFFFFu;
    for (int g = 0; g < 4; ++g) {
      if (head[g] >= (uint32_t)PL) continue;
      const uint32_t key = lk[(t * 4 + g) * PL + head[g]];
      if (key < kSkNoKey && key < bk) {
        best = g;
        bk = key;
      }
    }
    uint32_t id = kDummy;
    float v = FLT_MAX;
    if (best >= 0) {
      ++head[best];
      // Slot = point block * 4 + row; the lane's quarter of the block is g = best.
      const uint32_t slot = bk & smask;
      const uint32_t p = (slot >> 2) * 16 + best * 4 + (slot & 3u);
      id = a.perm[RBC_CHECKED(x0 + p, a.n)];
      v = -2.f * __uint_as_float(bk & vmask);
    }
    a.out_i[size_t(ent) * a.Kp + j] = id;
    a.out_v[size_t(ent) * a.Kp + j] = v;
  }
}

// ---- host: symmetric eigenvectors (cyclic Jacobi) -----------------------------------------
// A (n x n, row-major, symmetric) is diagonalised in place; V receives the eigenvectors as
// columns.  Rotations zero a_pq with t = sign(th) / (|th| + sqrt(th^2 + 1)), th = (a_qq -
// a_pp) / (2 a_pq); sweeps stop when the off-diagonal mass is below 1e-30 of the total.
void jacobi_eigen(std::vector<double> &A, uint32_t n, std::vector<double> &V) {
  V.assign(size_t(n) * n, 0.0);
  for (uint32_t i = 0; i < n; ++i) V[size_t(i) * n + i] = 1.0;
  double total = 0.0;
  for (double v : A) total += v * v;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0;
    for (uint32_t p = 0; p < n; ++p)
      for (uint32_t q = p + 1; q < n; ++q) off += 2.0 * A[size_t(p) * n + q] * A[size_t(p) * n + q];
    if (off <= 1e-30 * total) break;
    for (uint32_t p = 0; p + 1 < n; ++p)
      for (uint32_t q = p + 1; q < n; ++q) {
        const double apq = A[size_t(p) * n + q];
        if (fabs(apq) <= 1e-300) continue;
        const double th = (A[size_t(q) * n + q] - A[size_t(p) * n + p]) / (2.0 * apq);
        const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
        for (uint32_t k = 0; k < n; ++k) {  // columns p, q
          const double akp = A[size_t(k) * n + p], akq = A[size_t(k) * n + q];
          A[size_t(k) * n + p] = c * akp - s * akq;
          A[size_t(k) * n + q] = s * akp + c * akq;
        }
        for (uint32_t k = 0; k < n; ++k) {  // rows p, q
          const double apk = A[size_t(p) * n + k], aqk = A[size_t(q) * n + k];
          A[size_t(p) * n + k] = c * apk - s * aqk;
          A[size_t(q) * n + k] = s * apk + c * aqk;
        }
        for (uint32_t k = 0; k < n; ++k) {
          const double vkp = V[size_t(k) * n + p], vkq = V[size_t(k) * n + q];
          V[size_t(k) * n + p] = c * vkp - s * vkq;
          V[size_t(k) * n + q] = s * vkp + c * vkq;
        }
      }
  }
}

}  // namespace

int sketch_build(rbc_exact *h, hipStream_t st) {
  h->sk_dims = 0;
  h->sk_active = false;
  // Only where the sketch row (32 BF16) is shorter than the full one, for L2.
  if (h->metric != RBC_L2 || h->ldb < 64 || h->d > (uint32_t)kSkMaxD) return RBC_OK;
  const uint32_t n = h->n, d = h->d, ld = h->ld, R = h->R;
  const uint32_t S = std::min(n, kSkSample);
  DevBuf<double> mu64, part, C64, od2;
  const uint32_t nch = cdiv(S, kSkChunk);
  RBC_TRY(mu64.alloc(d));
  RBC_TRY(part.alloc(size_t(nch) * d * d));
  RBC_TRY(C64.alloc(size_t(d) * d));
  RBC_TRY(od2.alloc(S));
  // The mean's chunk partials use part's first nch * d entries before the covariance's.
  hipLaunchKernelGGL(sk_mean_part, dim3(cdiv(d, 64), nch), dim3(256), 0, st, h->xp.get(), (size_t)ld, n, S, d,
                     part.get());
  hipLaunchKernelGGL(sk_mean_sum, dim3(cdiv(d, 64)), dim3(64), 0, st, part.get(), nch, d, S, mu64.get());
  hipLaunchKernelGGL(sk_cov_part, dim3(cdiv(d, 32), cdiv(d, 32), nch), dim3(256), 0, st, h->xp.get(), (size_t)ld, n,
                     S, d, mu64.get(), part.get());
  hipLaunchKernelGGL(sk_cov_sum, dim3(cdiv(d * d, 256)), dim3(256), 0, st, part.get(), nch, d, S, C64.get());
  hipLaunchKernelGGL(sk_owner_d2, dim3(cdiv(S, 256)), dim3(256), 0, st, h->xp.get(), (size_t)ld, n, S, d,
                     h->perm.get(), h->owner.get(), h->reps.get(), od2.get());
  RBC_HIP_TRY(hipGetLastError());
  std::vector<double> mu(d), C(size_t(d) * d), o2(S);
  RBC_HIP_TRY(hipMemcpyAsync(mu.data(), mu64.get(), d * 8, hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipMemcpyAsync(C.data(), C64.get(), size_t(d) * d * 8, hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipMemcpyAsync(o2.data(), od2.get(), size_t(S) * 8, hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  double trace = 0.0, mo2 = 0.0;
  for (uint32_t i = 0; i < d; ++i) trace += C[size_t(i) * d + i];
  for (double v : o2) mo2 += v;
  mo2 /= (double)S;
  std::vector<double> V;
  jacobi_eigen(C, d, V);
  std::vector<uint32_t> order(d);
  for (uint32_t i = 0; i < d; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return C[size_t(a) * d + a] > C[size_t(b) * d + b]; });
  const uint32_t D = kSkDims;
  double kept = 0.0;
  for (uint32_t r = 0; r < D; ++r) kept += C[size_t(order[r]) * d + order[r]];
  // Rows of P: the D leading eigenvectors, re-orthonormalised (modified Gram-Schmidt,
  // twice) in FP64, then rounded to FP32.
  std::vector<double> P64(size_t(D) * d);
  for (uint32_t r = 0; r < D; ++r)
    for (uint32_t i = 0; i < d; ++i) P64[size_t(r) * d + i] = V[size_t(i) * d + order[r]];
  for (int pass = 0; pass < 2; ++pass)
    for (uint32_t r = 0; r < D; ++r) {
      double *pr = &P64[size_t(r) * d];
      for (uint32_t s = 0; s < r; ++s) {
        const double *ps = &P64[size_t(s) * d];
        double dot = 0.0;
        for (uint32_t i = 0; i < d; ++i) dot += pr[i] * ps[i];
        for (uint32_t i = 0; i < d; ++i) pr[i] -= dot * ps[i];
      }
      double nr = 0.0;
      for (uint32_t i = 0; i < d; ++i) nr += pr[i] * pr[i];
      nr = sqrt(nr);
      for (uint32_t i = 0; i < d; ++i) pr[i] /= nr;
    }
  std::vector<float> Pf(size_t(D) * ld, 0.f), muf(ld, 0.f);
  for (uint32_t r = 0; r < D; ++r)
    for (uint32_t i = 0; i < d; ++i) Pf[size_t(r) * ld + i] = (float)P64[size_t(r) * d + i];
  for (uint32_t i = 0; i < d; ++i) muf[i] = (float)mu[i];
  // |P_f|_2^2 = |P_f P_f^T|_2 <= 1 + |P_f P_f^T - I|_F; the FP64 products of FP32 values
  // are exact and the sums of d of them are within d 2^-53 of their absolute values.
  double g2 = 0.0;
  for (uint32_t r = 0; r < D; ++r)
    for (uint32_t s = 0; s < D; ++s) {
      double dot = 0.0;
      for (uint32_t i = 0; i < d; ++i) dot += (double)Pf[size_t(r) * ld + i] * (double)Pf[size_t(s) * ld + i];
      const double g = dot - (r == s ? 1.0 : 0.0);
      g2 += g * g;
    }
  h->sk_scale = sqrt(1.0 + sqrt(g2) + 1e-12 + d * 0x1p-52) * (1.0 + 1e-12);
  h->sk_resid = std::max(0.0, trace - kept);
  h->sk_total = trace;
  h->sk_owner_d2 = mo2;
  RBC_TRY(h->sk_P.alloc(size_t(D) * ld));
  RBC_TRY(h->sk_mu.alloc(ld));
  RBC_HIP_TRY(hipMemcpyAsync(h->sk_P.get(), Pf.data(), Pf.size() * 4, hipMemcpyHostToDevice, st));
  RBC_HIP_TRY(hipMemcpyAsync(h->sk_mu.get(), muf.data(), muf.size() * 4, hipMemcpyHostToDevice, st));
  // Sketch rows of the permuted database and their per-list maxima.
  RBC_TRY(h->xs.alloc(size_t(n) * kSkW));
  DevBuf<float> err, n2;
  RBC_TRY(err.alloc(n));
  RBC_TRY(n2.alloc(n));
  hipLaunchKernelGGL(sk_project, dim3(cdiv(n, kSkProjRows)), dim3(256), 0, st, h->xp.get(), (size_t)ld, n, d, h->sk_P.get(), ld,
                     h->sk_mu.get(), sketch_proj_err(d), 1, h->xs.get(), err.get(), n2.get());
  RBC_TRY(h->sk_lqe.alloc(R));
  RBC_TRY(h->sk_lnm.alloc(R));
  hipLaunchKernelGGL(sk_list_stats, dim3(R), dim3(256), 0, st, h->list_off.get(), err.get(), n2.get(),
                     h->sk_lqe.get(), h->sk_lnm.get());
  RBC_HIP_TRY(hipGetLastError());
  std::vector<float> lqe(R), lnm(R);
  RBC_HIP_TRY(hipMemcpyAsync(lqe.data(), h->sk_lqe.get(), size_t(R) * 4, hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipMemcpyAsync(lnm.data(), h->sk_lnm.get(), size_t(R) * 4, hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  h->sk_gqe = *std::max_element(lqe.begin(), lqe.end());
  h->sk_gnm = *std::max_element(lnm.begin(), lnm.end());
  h->sk_dims = D;
  sketch_set_mode(h, h->sk_mode);
  return RBC_OK;
}

double sketch_proj_err(uint32_t d) {
  const double u = 0x1p-24, N = (double)d + 1.0;
  return N * u / (1.0 - N * u) * 1.01;
}

void sketch_set_mode(rbc_exact *h, int mode) {
  h->sk_mode = mode;
  const bool worth = 2.0 * h->sk_resid <= kSkAutoFrac * h->sk_owner_d2;
  h->sk_active = h->sk_dims > 0 && (mode == RBC_SKETCH_ON || (mode == RBC_SKETCH_AUTO && worth));
}

int sketch_queries(const rbc_exact *h, const float *q, size_t ldq, uint32_t m, uint16_t *qs, float *qe, float *qn,
                   hipStream_t st) {
  RBC_ARG_CHECK(h->sk_dims > 0, "the index has no sketch (L2, 32 < d <= %d only)", kSkMaxD);
  if (m == 0) return RBC_OK;
  hipLaunchKernelGGL(sk_project, dim3(cdiv(m, kSkProjRows)), dim3(256), 0, st, q, ldq, m, h->d, h->sk_P.get(), h->ld,
                     h->sk_mu.get(), sketch_proj_err(h->d), 0, qs, qe, qn);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

double sketch_gemm_err() {
  const double up = 0x1p-23, K = (double)kSkW;
  return K * up / (1.0 - K * up) + 6.0 * up;
}

// Per query: the packed emission threshold (sk_threshold_pack) with the index-wide largest
// point error bound and sketch norm, which bound every list's.
__global__ void sk_query_thr(uint32_t m, const double *gam, const float *qe, const float *qn, double sP, float gqe,
                             float gnm, double c_gemm, uint32_t *thr) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < m) thr[q] = sk_threshold_pack(sP * gam[q], qe[q], qn[q], gqe, gnm, c_gemm);
}

// The scan's shape: 16-query blocks per wave, LDS candidate entries, next-fragment read,
// point blocks per step, ring stages, waves per workgroup.
constexpr int kSkQB = 8, kSkNW = (int)(kSkScanQ / (16 * kSkQB)), kSkCS = 1024, kSkPBStep = 8, kSkStagesN = 3;
static_assert(kSkNW * 16 * kSkQB == (int)kSkScanQ && kSkNW == 8, "scan items are cut at kSkScanQ queries");
using SkScan = void (*)(SkScanArgs);
constexpr SkScan kSkKernel = sk_scan_kernel<kSkQB, kSkCS, true, kSkPBStep, kSkStagesN, kSkNW>;

int sketch_probe(rbc_exact *h, const uint16_t *qs, const float *qn, const uint32_t *entries, uint32_t P,
                 const ScanItem *items, uint32_t nitems, uint32_t Kp, uint32_t *out_i, float *out_v, hipStream_t st) {
  if (nitems == 0) return RBC_OK;
  RBC_ARG_CHECK(Kp >= 1 && Kp <= 16, "sketch probe entries hold at most 16 candidates (K'=%u)", Kp);
  SkProbeArgs a;
  a.items = items;
  a.nitems = nitems;
  a.qids = entries;
  a.qs = qs;
  a.qn = qn;
  a.xs = h->xs.get();
  a.perm = h->perm.get();
  a.n = h->n;
  a.Kp = Kp;
  a.P = P;
  // Slot bits: 4 slots per 16-point block of the longest list.
  uint32_t max_len = 0;
  for (uint32_t r = 0; r < h->R; ++r) max_len = std::max(max_len, h->h_list_off[r + 1] - h->h_list_off[r]);
  const uint32_t slots = std::max(1u, (max_len + 15) / 16) * 4;
  uint32_t bits = 2;
  while ((1u << bits) < slots) ++bits;
  if (bits > 22) {
    set_error("sketch probe: a list of %u points is too long for its keys", max_len);
    return RBC_ERR_LIMIT;
  }
  a.idx_bits = bits;
  a.out_i = out_i;
  a.out_v = out_v;
  const dim3 grid(kXCD * ((nitems + kXCD - 1) / kXCD));
  if (Kp <= 8)
    hipLaunchKernelGGL((sk_probe_kernel<8, 3, 2>), grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((sk_probe_kernel<8, 3, 4>), grid, dim3(256), 0, st, a);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int sketch_scan(rbc_exact *h, uint32_t m, const uint16_t *qs, const float *qe, const float *qn, const double *gam,
                uint32_t *qthr, const uint32_t *qlist, const uint32_t *list_qoff, const ScanItem *items,
                uint32_t nitems, uint32_t *sched, uint32_t *cand_i, uint32_t *cand_n, uint32_t cap, hipStream_t st) {
  if (nitems == 0) return RBC_OK;
  hipLaunchKernelGGL(sk_query_thr, dim3(cdiv(m, 256)), dim3(256), 0, st, m, gam, qe, qn,
                     h->sk_scale * (1.0 + 1e-12), h->sk_gqe, h->sk_gnm, sketch_gemm_err(), qthr);
  hipLaunchKernelGGL((scan_split_kernel<kSkScanQ, 16 * kSkPBStep, 16 * kSkQB>), dim3(1), dim3(1024), 0, st, list_qoff,
                     h->list_len.get(), h->R, sched, sched + kXCD + 1);
  RBC_HIP_TRY(hipGetLastError());
  SkScanArgs a;
  a.items = items;
  a.bounds = sched;
  a.next = sched + kXCD + 1;
  a.qids = qlist;
  a.qthr = qthr;
  a.qs = qs;
  a.xs = h->xs.get();
  a.perm = h->perm.get();
  a.n = h->n;
  a.cand_i = cand_i;
  a.cand_n = cand_n;
  a.cap = cap;
  // One persistent block per resident slot: (CUs x blocks per CU), a multiple of kXCD so
  // that every XCD gets blocks.
  static int per_cu = 0, cus = 0;
  if (per_cu == 0) {
    int dev = 0;
    RBC_HIP_TRY(hipGetDevice(&dev));
    RBC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RBC_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kSkKernel, 64 * kSkNW, 0));
    per_cu = std::max(per_cu, 1);
  }
  uint32_t grid = (uint32_t)(cus * per_cu);
  grid = std::min(grid, nitems + kXCD);
  grid = cdiv(std::max(grid, (uint32_t)kXCD), kXCD) * kXCD;
  hipLaunchKernelGGL(kSkKernel, dim3(grid), dim3(64 * kSkNW), 0, st, a);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

}  // namespace rbc

using namespace rbc;

extern "C" {

int rbc_exact_set_sketch(rbc_exact *h, int mode) {
  RBC_ARG_CHECK(h, "NULL handle");
  RBC_ARG_CHECK(mode == RBC_SKETCH_AUTO || mode == RBC_SKETCH_OFF || mode == RBC_SKETCH_ON, "unknown sketch mode %d",
                mode);
  RBC_ARG_CHECK(mode != RBC_SKETCH_ON || h->sk_dims > 0, "the index has no sketch (L2, 32 < d <= %d only)", kSkMaxD);
  sketch_set_mode(h, mode);
  return RBC_OK;
}

int rbc_exact_sketch_info(const rbc_exact *h, rbc_sketch_info *info) {
  RBC_ARG_CHECK(h && info, "NULL argument");
  info->dims = h->sk_dims;
  info->active = h->sk_active ? 1u : 0u;
  info->norm_bound = h->sk_scale;
  info->resid_var = h->sk_resid;
  info->total_var = h->sk_total;
  info->mean_owner_d2 = h->sk_owner_d2;
  return RBC_OK;
}

int rbc_exact_sketch_rows(rbc_exact *h, const float *x, uint32_t m, uint32_t ldx, int mem, uint16_t *out, float *err,
                          float *n2) {
  RBC_ARG_CHECK(h && (x || m == 0), "NULL argument");
  RBC_ARG_CHECK(h->sk_dims > 0, "the index has no sketch (L2, 32 < d <= %d only)", kSkMaxD);
  RBC_ARG_CHECK(ldx >= h->d, "ldx=%u below d=%u", ldx, h->d);
  if (m == 0) return RBC_OK;
  hipStream_t st = nullptr;
  DevMat X;
  RBC_TRY(X.stage(x, m, h->d, ldx, mem, st));
  DevBuf<uint16_t> o;
  DevBuf<float> e, n;
  RBC_TRY(o.alloc(size_t(m) * kSkW));
  RBC_TRY(e.alloc(m));
  RBC_TRY(n.alloc(m));
  RBC_TRY(sketch_queries(h, X.p, X.ld, m, o.get(), e.get(), n.get(), st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  if (out) RBC_TRY(copy_out(out, o.get(), size_t(m) * kSkW * 2, mem, st));
  if (err) RBC_TRY(copy_out(err, e.get(), size_t(m) * 4, mem, st));
  if (n2) RBC_TRY(copy_out(n2, n.get(), size_t(m) * 4, mem, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int rbc_exact_sketch_basis(const rbc_exact *h, float *P, float *mu) {
  RBC_ARG_CHECK(h, "NULL handle");
  RBC_ARG_CHECK(h->sk_dims > 0, "the index has no sketch (L2, 32 < d <= %d only)", kSkMaxD);
  if (P)
    RBC_HIP_TRY(hipMemcpy2D(P, size_t(h->d) * 4, h->sk_P.get(), size_t(h->ld) * 4, size_t(h->d) * 4, h->sk_dims,
                            hipMemcpyDeviceToHost));
  if (mu) RBC_HIP_TRY(hipMemcpy(mu, h->sk_mu.get(), size_t(h->d) * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

}  // extern "C"
