// BF16 MFMA distance filter of the exact query (precision RBC_PREC_BF16).
//
// The list scan evaluates q~.x~ on BF16 copies (q~ = bf16(q), x~ = bf16(x)) with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation) and emits every point whose TRUE
// distance could still be within G, the probe's upper bound on the k-th NN distance.
// With eta = |q - q~| + max_list |x - x~| (FP64, rounded up), |q - x| <= G implies
// |q~ - x~| <= G + eta, which the kernel tests as -|x~|^2/2 + q~.x~ >= -s: the
// accumulator starts at -|x~|^2/2 and s carries a margin of (gamma_K + 6u') times the
// absolute terms (u' = 2^-23, so no rounding mode of the accumulator is assumed)
// (bf16_threshold).  The test is conservative; the exact FP32 re-rank decides.
//
// Scan work item = (one ownership list, up to 512 queries that need it).  The four waves
// of a workgroup keep 128 queries each as MFMA B operands in VGPRs (two waves per SIMD)
// and share the list's points, the A operands, 32 per step, through a four-stage LDS
// ring filled three steps ahead by direct global->LDS loads.  Each ring read feeds 8 MFMAs,
// and each lane's accumulators hold four points of ONE query per query block, so the
// per-query threshold is one register per block.  The grid is persistent; the list-major
// items are split into one contiguous range per XCD (of equal work), so the items of a
// list run on one XCD at the same time and re-read the list from that XCD's L2.  The probe
// (close points of one list per (query, list) entry) uses 64-entry items on a three-stage
// register-staged LDS ring, each lane keeping its best few packed (value, slot) keys in
// registers.
#include "exact_query.hpp"
#include "scan_sched.hpp"

namespace rbc {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// One wave per query: BF16 copy (zero padded to ldb), |q~|^2 and |q - q~| (FP64); with
// qlo, also the low halves bf16(q - q~) of the split operands (the FP32 difference is
// exact); with qn2, also ||q||^2 (FP64 sum rounded to nearest) and the absolute error bound
// of the split-BF16 query-rep distances of this query (qr_error).
__global__ __launch_bounds__(256) void prep_queries(const float *q, size_t ldq, uint32_t m, uint32_t d,
                                                    uint32_t ldb, uint16_t *qb, uint16_t *qlo, float *qn, float *qe,
                                                    float *qn2, float *qerr, double c_split, double rep_nmax) {
  const uint32_t r = blockIdx.x * 4 + threadIdx.x / kWave;
  if (r >= m) return;
  const int l = lane_id();
  double nb = 0.0, e2 = 0.0, n2 = 0.0;
  for (uint32_t c = l; c < ldb; c += kWave) {
    float v = c < d ? q[size_t(r) * ldq + c] : 0.f;
    uint16_t h = f32_to_bf16_rne(v);
    qb[size_t(r) * ldb + c] = h;
    if (qlo) qlo[size_t(r) * ldb + c] = f32_to_bf16_rne(v - __uint_as_float(uint32_t(h) << 16));
    double hv = (double)__uint_as_float(uint32_t(h) << 16);
    nb += hv * hv;
    double e = (double)v - hv;
    e2 += e * e;
    n2 += (double)v * (double)v;
  }
  for (int s = kWave / 2; s > 0; s >>= 1) {
    nb += __shfl_xor(nb, s, kWave);
    e2 += __shfl_xor(e2, s, kWave);
    n2 += __shfl_xor(n2, s, kWave);
  }
  if (l == 0) {
    qn[r] = (float)nb;
    float ef = (float)sqrt(e2);
    if ((double)ef < sqrt(e2)) ef = nextafterf(ef, FLT_MAX);
    qe[r] = ef;
    if (qn2) {
      qn2[r] = (float)n2;
      // |D~ - D| for every rep r (qr_mfma): the dot product's error 2 c |q| |r|, the
      // rounding of the two stored norms (u each) and of the epilogue's add and FMA
      // (u (|q|^2 + |r|^2) and u (|q|^2 + |r|^2 + 2 |q.r~|)), with |r| <= rep_nmax.
      const double u = 0x1p-24, nq = sqrt(n2), nr = rep_nmax;
      const double E = (2.0 * c_split * nq * nr + 4.0 * u * (n2 + nr * nr) + 2.0 * u * (1.0 + c_split) * nq * nr) *
                           1.01 + 1e-30;
      float Ef = (float)E;
      if ((double)Ef < E) Ef = nextafterf(Ef, FLT_MAX);
      qerr[r] = Ef;
    }
  }
}

// ---- query -> rep distances on MFMA (L2) --------------------------------------------
// D~[q][r] = |q|^2 + |r|^2 - 2 q.r with q.r from split BF16 operands: q = qh + ql + dq,
// r = rh + rl + dr (hi = bf16(v), lo = bf16(v - hi), |d| <= 2^-16 |v| per coordinate), and
// q.r ~ qh.rh + qh.rl + ql.rh on three v_mfma_f32_16x16x32_bf16 per tile and k-step.  The
// dropped terms are at most 3.1 2^-16 sum |q_i r_i|, the FP32 accumulation of the 3 K
// exact products at most gamma_{3K+1} (u' = 2^-23) of their absolute sum (<= 1.02 sum
// |q_i r_i|): together c_split |q| |r| (qr_split_err).  A = 16 reps, B = 16 queries, so
// lane l holds reps 4 (l/16) .. +3 of query l % 16 and writes them as one 16-byte run of
// D's row.  A wave computes 64 reps x 64 queries, a block of four waves 128 x 128; the
// 16 rep tiles of a query tile run on one XCD (block b on XCD b % 8), so the query rows
// are fetched from HBM once.  The queries' hi/lo rows come split by prep_queries and the
// reps' (R x ldb, zero padded) stay in L2: every operand is one 16-byte load.
constexpr int kQM = 128;  // queries and reps per block

template <int KS>
__global__ __launch_bounds__(256) void qr_mfma(const uint16_t *qhi, const uint16_t *qlo, uint32_t m,
                                               const uint16_t *rhi, const uint16_t *rlo, uint32_t ldb,
                                               const float *rn2, uint32_t R, const float *qn2, float *D,
                                               uint32_t nqt, uint32_t nrt) {
  const uint32_t b = blockIdx.x, xcd = b % kXCD, s = b / kXCD;
  const uint32_t rt = s % nrt, qt = (s / nrt) * kXCD + xcd;
  if (qt >= nqt) return;
  const int w = threadIdx.x / kWave, l = lane_id();
  const uint32_t q0 = qt * kQM + (w / 2) * 64, r0 = rt * kQM + (w % 2) * 64;
  const uint32_t koff = (l / 16) * 8;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const uint32_t c0 = ks * 32 + koff;
    bf16x8 Ah[4], Al[4], Bh[4], Bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t r = r0 + i * 16 + (l % 16);
      if (r < R) {
        Ah[i] = *reinterpret_cast<const bf16x8 *>(rhi + size_t(r) * ldb + c0);
        Al[i] = *reinterpret_cast<const bf16x8 *>(rlo + size_t(r) * ldb + c0);
      } else {
        Ah[i] = bf16x8{};
        Al[i] = bf16x8{};
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t q = q0 + j * 16 + (l % 16);
      if (q < m) {
        Bh[j] = *reinterpret_cast<const bf16x8 *>(qhi + size_t(q) * ldb + c0);
        Bl[j] = *reinterpret_cast<const bf16x8 *>(qlo + size_t(q) * ldb + c0);
      } else {
        Bh[j] = bf16x8{};
        Bl[j] = bf16x8{};
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[i], Bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[i], Bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al[i], Bh[j], acc[i][j], 0, 0, 0);
      }
  }
  // Epilogue: lane l holds reps r0 + 16 i + 4 (l/16) + e of query q0 + 16 j + l % 16.
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t q = q0 + j * 16 + (l % 16);
    if (q >= m) continue;
    const float nq = qn2[q];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t rb = r0 + i * 16 + (l / 16) * 4;
      float o[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float nr = rb + e < R ? rn2[rb + e] : 0.f;
        o[e] = fmaxf(__builtin_fmaf(-2.f, acc[i][j][e], nq + nr), 0.f);
      }
      float *dst = D + size_t(q) * R + rb;
      if (rb + 3 < R && (R % 4) == 0) {
        *reinterpret_cast<float4 *>(dst) = make_float4(o[0], o[1], o[2], o[3]);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (rb + e < R) dst[e] = o[e];
      }
    }
  }
}

struct Bf16ScanArgs {
  const ScanItem *items;
  const uint32_t *bounds;  // [kXCD + 1]: items [bounds[x], bounds[x+1]) run on XCD x
  uint32_t *next;          // [kXCD]: items of XCD x's range taken so far
  const uint32_t *qids;  // list-major (query, list) pairs: query ids
  const uint16_t *qb;    // queries, BF16, ldb
  const float *qn;       // |q~|^2
  const float *qe;       // |q - q~|
  const double *gam;     // per query: upper bound on the k-th NN distance (metric units)
  const float *lqe;      // per list: max |x - x~|
  const float *lnm;      // per list: max |x~|^2
  double c_gemm;         // gamma_K + 6u' (u' = 2^-23)
  const uint16_t *xb;    // permuted database, BF16, ldb
  const float *xn;       // |x~|^2 per permuted row
  const uint32_t *perm;  // permuted row -> original id
  uint32_t n;            // database rows (bounds checks)
  uint32_t ldb;
  float *cand_d;
  uint32_t *cand_i;
  uint32_t *cand_n;
  uint32_t cap;
};

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

// ---- BF16 list scan ------------------------------------------------------------------
// Work item = (one list, up to 512 queries).  Wave w keeps queries 128w..128w+127 as MFMA
// B operands in VGPRs (8 blocks of 16); the 32 points of a step, the A operands (two
// blocks of 16), are shared by all four waves through a four-stage LDS ring that is filled by direct global->LDS loads (global_load_lds_dwordx4)
// three steps ahead, so no VGPRs hold data in flight.  Each wave waits for its own loads of
// the step (s_waitcnt vmcnt with the count of younger loads) before the step's barrier.
// A stage holds 32 rows of 256 bytes; the 16-byte chunk c of row r sits at chunk c ^ (r%16),
// which keeps the 64-lane reads (16 rows x 4 chunks) free of bank conflicts.  A ring read
// (16 points x 32 coordinates) feeds 8 MFMAs, one per query block: twice as many as with 64
// queries per wave, which halves the LDS reads, ring loads and barriers per FLOP.  The
// output tile of a (point block, query block) MFMA gives lane l query l%16 and points
// 4 (l/16) .. +3, so the accumulators start at those points' -|x~|^2/2 (one 16-byte LDS
// read of the ring's norms) and the per-query threshold is one register.  The LDS DMA
// is issued from asm: a compiler-visible one would make the compiler drain it before
// every LDS read.
// Candidates go to an LDS list first (an LDS atomic does not wait for the loads in flight)
// and are appended to the per-query buffers after the last step; entries beyond the LDS
// list's capacity are appended directly.
constexpr int kQB = 8;           // 16-query blocks per wave (128 queries)
constexpr int kVQ = 64 * kQB;    // queries per item (4 waves)
static_assert(kVQ == (int)kBf16ScanQ, "scan items are cut at kBf16ScanQ queries");
constexpr int kVP = 32;          // points per step
constexpr int kSStages = 4;      // scan ring stages (prefetch distance 3)
constexpr int kSStageB = kVP * 256;
constexpr int kCS = 1024;        // LDS candidate list entries per item
// Probe ring (register-staged): rows padded to 288 bytes, which measured free of bank
// conflicts for the same 64-lane read pattern.
constexpr int kRowB = 288;
constexpr int kStages = 3;

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(3))) const float lds_float;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4;
typedef __attribute__((address_space(3))) const i32x4 lds_i32x4;

__device__ __forceinline__ int max3i(int a, int b, int c) { return max(max(a, b), c); }

__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(const lds_void *)p;
}

// Direct global->LDS loads (global_load_lds_dword[x4]: lane l writes bytes [16 l, 16 l + 16)
// or [4 l, 4 l + 4) after the wave-uniform LDS address in M0).  Issued from asm so that the compiler's wait insertion,
// which would otherwise drain them before unrelated instructions, leaves them alone: the
// kernel waits for them itself (wait_vm) before the barrier that publishes the data.
// M0 is saved and restored inside the block, so the compiler's use of it is unaffected;
// the s_nop is the wait state the hardware needs between writing M0 and an LDS DMA that
// reads it (the compiler inserts the same for its own LDS DMA).
// The three ring loads of one wave for one step: two 1 KB row blocks and the norms, from
// scalar base addresses (xs, ns) with 32-bit per-lane byte offsets, into LDS at m0a, m0b
// and m0n.
__device__ __forceinline__ void dma_step(const void *xs, const void *ns, uint32_t o0, uint32_t o1, uint32_t on,
                                         uint32_t m0a, uint32_t m0b, uint32_t m0n) {
  uint32_t keep;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %4\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %1, %7\n\t"
      "s_mov_b32 m0, %5\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %2, %7\n\t"
      "s_mov_b32 m0, %6\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dword %3, %8\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(keep)
      : "v"(o0), "v"(o1), "v"(on), "s"(m0a), "s"(m0b), "s"(m0n), "s"(xs), "s"(ns)
      : "memory");
}

// vmcnt(n) with the other counters left alone (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

template <int KS>
__global__ __launch_bounds__(256, 2) void bf16_scan_kernel(Bf16ScanArgs a) {
  // Emission test per (query i, point x): MFMA(s_i + q~.x~) >= |x~|^2 / 2 with s_i the
  // threshold of the (query, list) pair (bf16_threshold), so the test never drops an x with
  // |q - x| <= G_i.
  __shared__ float sh[kVQ];
  __shared__ uint32_t qrow[kVQ];
  __shared__ __attribute__((aligned(16))) unsigned char ring[kSStages * kSStageB];
  __shared__ __attribute__((aligned(16))) float nring[kSStages][4][kWave];  // ||x~||^2, per wave copy
  __shared__ uint32_t cs_p[kCS];  // candidate: point index in the list
  __shared__ float cs_v[kCS];     // candidate: approximate squared distance minus |q~|^2
  __shared__ uint16_t cs_q[kCS];   // candidate: query slot in the item
  __shared__ uint32_t cs_n;
  __shared__ uint32_t s_item;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  // Persistent blocks: block b serves XCD b % kXCD (workgroups are dispatched to the XCDs
  // round robin) and takes that XCD's items in order, so the items of a list run at the
  // same time on one XCD and re-read the list from its L2.
  // A block whose XCD range is exhausted takes items from the other ranges (in XCD order
  // after its own), so the XCDs finish together.
  const uint32_t xcd = blockIdx.x % kXCD;
  const uint32_t lo = a.bounds[xcd], hi = a.bounds[xcd + 1];
  // LDS destinations of this wave's ring loads (stage 0).
  const uint32_t wu = __builtin_amdgcn_readfirstlane((uint32_t)w);
  const uint32_t m0x = lds_addr(ring) + wu * 2048, m0n = lds_addr(&nring[0][0][0]) + wu * (kWave * 4);
  // Per-lane parts of the ring loads: block j of this wave holds rows 8w + 4j + l/16, and
  // lane l fills chunk slot l%16 of its row with the chunk that belongs there (chunks past
  // the row length load chunk 0).
  uint32_t lrow[2], lcb[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    lrow[j] = w * 8 + j * 4 + l / 16;
    const uint32_t lc = (uint32_t)(l % 16) ^ (lrow[j] & 15);
    lcb[j] = (lc < (uint32_t)KS * 4 ? lc : 0) * 16;
  }
  for (;;) {
  if (t == 0) {
    uint32_t v = lo + atomicAdd(&a.next[xcd], 1u);
    if (v >= hi) {
      v = kDummy;
      for (uint32_t d = 1; d < (uint32_t)kXCD; ++d) {
        const uint32_t y = (xcd + d) % kXCD;
        const uint32_t u = a.bounds[y] + atomicAdd(&a.next[y], 1u);
        if (u < a.bounds[y + 1]) {
          v = u;
          break;
        }
      }
    }
    s_item = v;
  }
  __syncthreads();
  const uint32_t item = s_item;
  if (item == kDummy) break;
  const ScanItem it = a.items[item];
  const uint32_t x0 = __builtin_amdgcn_readfirstlane(it.x_begin), np = __builtin_amdgcn_readfirstlane(it.x_count);
  const uint32_t nsteps = (np + kVP - 1) / kVP;
  // Loads of one step: two 1 KB blocks per wave (rows 8w..8w+7, all 16 chunks; lane l
  // fills chunk slot l%16 of row 8w + 4j + l/16 with the chunk that belongs there) and the
  // norms of the step's 32 points (lanes 32..63 repeat lanes 0..31).  Rows past the list end
  // and chunks past the row length load clamped, valid addresses: every lane issues its
  // load, so each wave issues exactly three loads per step.
  auto issue = [&](uint32_t step) {
    const uint32_t stage = step % kSStages, p0 = step * kVP, last = np - 1 - p0;
    const uint16_t *xs = a.xb + size_t(x0 + p0) * a.ldb;
    const float *ns = a.xn + x0 + p0;
    const uint32_t m0a = m0x + stage * kSStageB;
    dma_step(xs, ns, min(lrow[0], last) * (a.ldb * 2) + lcb[0], min(lrow[1], last) * (a.ldb * 2) + lcb[1],
             min((uint32_t)(l & 31), last) * 4, m0a, m0a + 1024, m0n + stage * (4 * kWave * 4));
  };
  // The ring's first loads depend only on the item, so they go out before the query setup.
  if (nsteps > 0) issue(0);
  if (nsteps > 1) issue(1);
  if (nsteps > 2) issue(2);
  // Wave w owns query slots 128 w .. 128 w + 127 and lane l fetches the ids of slots
  // 128 w + 64 h + l (h = 0, 1), so a wave already holds the rows of its own queries: the
  // MFMA operands are fetched through lane shuffles, in parallel with the threshold
  // inputs, without waiting for an LDS copy of the ids.
  uint32_t myq[kQB / 4];
#pragma unroll
  for (int h = 0; h < kQB / 4; ++h) {
    const uint32_t slot = 16 * kQB * w + 64 * h + l;
    myq[h] = slot < it.q_count ? a.qids[it.q_begin + slot] : kDummy;
  }
  const bool wave_live = (uint32_t)(w * 16 * kQB) < it.q_count;
  bf16x8 A[kQB][KS];
  const uint32_t koff = (l / 16) * 8;
#pragma unroll
  for (int qb = 0; qb < kQB; ++qb) {
    const uint32_t row = (uint32_t)__shfl((int)myq[qb / 4], (qb % 4) * 16 + (l % 16), kWave);
#pragma unroll
    for (int s = 0; s < KS; ++s)
      A[qb][s] = row != kDummy ? load8(a.qb + size_t(row) * a.ldb + s * 32 + koff) : bf16x8{};
  }
#pragma unroll
  for (int h = 0; h < kQB / 4; ++h) {
    const uint32_t slot = 16 * kQB * w + 64 * h + l;
    qrow[slot] = myq[h];
    sh[slot] = myq[h] != kDummy ? bf16_threshold(a.gam[myq[h]], a.qe[myq[h]], a.qn[myq[h]], a.lqe[it.aux],
                                                  a.lnm[it.aux], a.c_gemm)
                                : -INFINITY;
  }
  if (t == 0) cs_n = 0;
  __syncthreads();
  // Lane l's query of block qb is slot 128 w + 16 qb + l%16; it emits x when the
  // accumulator reaches -s.  The pre-filter compares float bit patterns as signed integers,
  // which orders them like the floats for a positive threshold; a threshold <= 0 lets
  // every value through to the exact float test.
  int thrI[kQB];
#pragma unroll
  for (int qb = 0; qb < kQB; ++qb) {
    const float tq = -sh[w * 16 * kQB + qb * 16 + (l % 16)];
    thrI[qb] = tq > 0.f ? __float_as_int(tq) : INT_MIN;
  }
  // Per-lane read addresses (stage 0): point fragment (pb, ks) is row pb*16 + l%16, chunk
  // l/16 + 4 ks; rows 16..31 sit 4096 bytes after rows 0..15 with the same swizzle (r%16 is
  // the same), so four addresses serve both point blocks.  Chunks at or past 4 KS are read
  // but never used.
  uint32_t ad0[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const uint32_t r = l % 16;
    const uint32_t c = ((uint32_t)(l / 16) + 4 * ks) ^ r;
    ad0[ks] = lds_addr(ring) + r * 256 + c * 16;
  }
  // The query operands must have arrived before the first MFMA; waiting here (for everything,
  // the ring's first loads included) keeps the compiler from placing that wait inside the
  // loop, where it would also drain the ring loads in flight.
  wait_vm<0>();
  for (uint32_t s = 0; s < nsteps; ++s) {
    // This wave's loads of step s are done when at most the loads of steps s+1 and s+2
    // (those that exist) are outstanding; the barrier then publishes every wave's.
    if (s + 2 < nsteps)
      wait_vm<6>();
    else if (s + 1 < nsteps)
      wait_vm<3>();
    else
      wait_vm<0>();
    asm volatile("s_barrier" ::: "memory");
    if (s + 3 < nsteps) issue(s + 3);
    if (!wave_live) continue;
    const uint32_t so = (s % kSStages) * kSStageB;
    // Plain LDS reads: the compiler does not see the ring's LDS DMA (it is issued from
    // asm), so it neither drains those loads before the reads nor moves the reads above
    // the barrier (whose asm clobbers memory); it waits for the reads with lgkmcnt.
    // Rows past the list end hold copies of its last row; the exact test drops them.
#pragma unroll
    for (int pb = 0; pb < 2; ++pb) {
      i32x4 braw[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        braw[i] = *reinterpret_cast<const lds_i32x4 *>((uintptr_t)(ad0[i] + so + pb * 4096));
      // -|x~|^2/2 of the chunk's rows (l/16)*4 .. +3 as the initial accumulator.
      const float4 n4 = *reinterpret_cast<const float4 *>(&nring[s % kSStages][w][pb * 16 + (l / 16) * 4]);
      const f32x4 init = f32x4{-0.5f * n4.x, -0.5f * n4.y, -0.5f * n4.z, -0.5f * n4.w};
      f32x4 acc[kQB];
      // While a wave issues its MFMAs it goes first: the SIMD's other wave fills the gaps
      // between its MFMAs instead of delaying them.
      __builtin_amdgcn_s_setprio(1);
      {
        bf16x8 B[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) B[ks] = __builtin_bit_cast(bf16x8, braw[ks]);
#pragma unroll
        for (int qb = 0; qb < kQB; ++qb) acc[qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[0], A[qb][0], init, 0, 0, 0);
#pragma unroll
        for (int ks = 1; ks < KS; ++ks)
#pragma unroll
          for (int qb = 0; qb < kQB; ++qb)
            acc[qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[ks], A[qb][ks], acc[qb], 0, 0, 0);
      }
      __builtin_amdgcn_s_setprio(0);
      // One pre-filter flag per query block, so that a chunk with a hit only tests the
      // blocks that have one.
      bool g[kQB];
      bool hit = false;
#pragma unroll
      for (int qb = 0; qb < kQB; ++qb) {
        g[qb] = max(max3i(__float_as_int(acc[qb][0]), __float_as_int(acc[qb][1]), __float_as_int(acc[qb][2])),
                    __float_as_int(acc[qb][3])) >= thrI[qb];
        hit = hit || g[qb];
      }
      if (__any(hit)) {
        const uint32_t p0 = s * kVP + pb * 16 + (l / 16) * 4;
#pragma unroll
        for (int qb = 0; qb < kQB; ++qb) {
          if (!g[qb]) continue;
          const int qi = w * 16 * kQB + qb * 16 + (l % 16);
          const float th = -sh[qi];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const uint32_t p = p0 + r;
            if (acc[qb][r] >= th && p < np) {
              const float v = -2.f * acc[qb][r];
              const uint32_t slot = atomicAdd(&cs_n, 1u);
              if (slot < (uint32_t)kCS) {
                cs_p[slot] = p;
                cs_v[slot] = v;
                cs_q[slot] = (uint16_t)qi;
              } else {
                const uint32_t row = qrow[qi];
                const uint32_t gs = atomicAdd(&a.cand_n[row], 1u);
                if (gs < a.cap) {
                  a.cand_d[size_t(row) * a.cap + gs] = a.qn[row] + v;
                  a.cand_i[size_t(row) * a.cap + gs] = a.perm[RBC_CHECKED(x0 + p, a.n)];
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
      a.cand_d[size_t(row) * a.cap + gs] = a.qn[row] + cs_v[i];
      a.cand_i[size_t(row) * a.cap + gs] = a.perm[RBC_CHECKED(x0 + cs_p[i], a.n)];
    }
  }
  // The next item rewrites qrow, sh, cs_* and s_item.
  __syncthreads();
  }
}

// ---- BF16 probe -----------------------------------------------------------------------
// Work item = (one list, up to 64 probe entries; entry e belongs to query e / P).  The MFMA
// operands are swapped with respect to the scan: A = 16 points of the list (from a
// three-stage LDS ring), B = 16 queries of the wave, so that lane l's accumulator holds
// four points of ONE entry (column l%16).  The four lanes l, l^16, l^32, l^48 see disjoint
// quarters of the list for their entry; each keeps its kPL smallest keys sorted in
// registers.  A key is the approximate distance's float bits with the low idx_bits bits
// replaced by the point's slot in the lane's quarter (step, tile, row), so keys are unique
// per lane, order by value first (to 23 - idx_bits mantissa bits) and then by slot, and a
// key is inserted branch-free with one v_med3_u32 per kept slot.  After the scan the four
// lists of an entry are merged by one thread into its min(K', 4 kPL) best (the remaining
// of its K' slots are empty).  The caller merges the entries of a query, re-ranks the
// best K' exactly and takes the k-th: any k distinct points bound the k-th NN distance
// from above, so the probe only has to find close points, not the exact top-K' of each
// list.  A query's k nearest neighbours are spread over several lists, and 4 per lane (16
// per entry) scan 0.03 % more pairs than 8 per lane at a quarter less probe time.
constexpr int kPL = 4;                 // kept per lane
constexpr int kPK = 32;                // largest K'
static_assert(4 * kPL == (int)kBf16ProbeKeep, "probe entries hold 4 lanes x kPL keys");
constexpr int kPQ = kBf16ProbeQ;       // queries per probe item
constexpr uint32_t kKeyInf = 0x7F800000u;  // keys at or above this hold no point

// Median of three unsigned words (one VALU instruction).
__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// Inserts key into the ascending list t (the largest drops out): slot j takes the median of
// (t[j-1], key, t[j]), which is t[j-1], key or t[j] as key falls below, between or above.
__device__ __forceinline__ void insert_key(uint32_t (&t)[kPL], uint32_t key) {
#pragma unroll
  for (int j = kPL - 1; j > 0; --j) t[j] = med3_u32(t[j - 1], key, t[j]);
  t[0] = min(t[0], key);
}

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
  uint32_t P;        // entries per query: entry e belongs to query e / P
  uint32_t idx_bits; // low key bits holding a point's slot (8 slots per 32-point step)
  uint32_t *out_i;   // [entries][Kp] original ids, ascending approximate distance
  float *out_v;      // [entries][Kp] their approximate squared distances D~ (truncated)
};

template <int KS>
__global__ __launch_bounds__(256, 2) void bf16_probe_kernel(Bf16ProbeArgs a) {
  // Workgroup b runs on XCD b % kXCD; each XCD takes a contiguous run of the list-major
  // items, so the items of one list re-read it from the same L2.
  const uint32_t per_xcd = (a.nitems + kXCD - 1) / kXCD;
  const uint32_t item = (blockIdx.x % kXCD) * per_xcd + blockIdx.x / kXCD;
  if (item >= a.nitems) return;
  const ScanItem it = a.items[item];
  __shared__ uint32_t qrow[kPQ];
  __shared__ __attribute__((aligned(16))) unsigned char ring[kStages][kVP * kRowB];
  __shared__ __attribute__((aligned(16))) float xring[kStages][kVP];  // ||x~||^2 (INF past the end)
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  if (t < kPQ) qrow[t] = (uint32_t)t < it.q_count ? a.qids[it.q_begin + t] : kDummy;
  __syncthreads();
  const uint32_t myent = qrow[w * 16 + (l % 16)];
  const uint32_t myrow = myent != kDummy ? myent / a.P : kDummy;
  const bool wave_live = (uint32_t)(w * 16) < it.q_count;
  const uint32_t koff = (l / 16) * 8;
  bf16x8 Bq[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s)
    Bq[s] = myrow != kDummy ? load8(a.qb + size_t(myrow) * a.ldb + s * 32 + koff) : bf16x8{};
  const float qn = myrow != kDummy ? a.qn[myrow] : 0.f;
  const uint32_t vmask = ~((1u << a.idx_bits) - 1u);
  uint32_t tk[kPL];
#pragma unroll
  for (int j = 0; j < kPL; ++j) tk[j] = 0xFFFFFFFFu;
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
        // Negative values (rounding) clamp to 0; padded points (xn = INF) give INF keys.
        const float v = fmaxf(__builtin_fmaf(-2.f, acc[r], xv[r]) + qn, 0.f);
        insert_key(tk, (__float_as_uint(v) & vmask) | (s * 8 + pt * 4 + r));
      }
    }
  }
  // Hand the per-lane lists to one thread per query through LDS (the ring is free now).
  __syncthreads();
  uint32_t *lk = reinterpret_cast<uint32_t *>(&ring[0][0]);        // [kPQ][4][kPL]
  if (wave_live) {
    const int q = w * 16 + (l % 16), g = l / 16;
#pragma unroll
    for (int j = 0; j < kPL; ++j) lk[(q * 4 + g) * kPL + j] = tk[j];
  }
  __syncthreads();
  if (t >= kPQ || qrow[t] == kDummy) return;
  const uint32_t row = qrow[t];  // the entry
  const uint32_t smask = ~vmask;
  uint32_t head[4] = {0, 0, 0, 0};
  for (uint32_t j = 0; j < a.Kp; ++j) {
    int best = -1;
    uint32_t bk = 0xFFFFFFFFu;
    for (int g = 0; g < 4; ++g) {
      if (head[g] >= (uint32_t)kPL) continue;
      const uint32_t key = lk[(t * 4 + g) * kPL + head[g]];
      if (key < kKeyInf && key < bk) {
        best = g;
        bk = key;
      }
    }
    uint32_t id = kDummy;
    float v = FLT_MAX;
    if (best >= 0) {
      ++head[best];
      // Slot = step * 8 + tile * 4 + row; the lane's quarter is g = best.
      const uint32_t slot = bk & smask;
      const uint32_t p = (slot >> 3) * kVP + ((slot >> 2) & 1u) * 16 + best * 4 + (slot & 3u);
      id = a.perm[x0 + p];
      v = __uint_as_float(bk & vmask);
    }
    a.out_i[size_t(row) * a.Kp + j] = id;
    a.out_v[size_t(row) * a.Kp + j] = v;
  }
}

template <int KS>
void launch_probe(const Bf16ProbeArgs &a, uint32_t nitems, hipStream_t st) {
  hipLaunchKernelGGL((bf16_probe_kernel<KS>), dim3(kXCD * ((nitems + kXCD - 1) / kXCD)), dim3(256), 0, st, a);
}

}  // namespace

int bf16_prepare_queries(rbc_exact *h, const float *q, size_t ldq, uint32_t m, uint16_t *qb, float *qn,
                         float *qe, hipStream_t st, float *qn2, float *qerr, uint16_t *qlo) {
  hipLaunchKernelGGL(prep_queries, dim3(cdiv(m, 4)), dim3(256), 0, st, q, ldq, m, h->d, h->ldb, qb, qlo, qn, qe, qn2,
                     qerr, qr_split_err(h), h->rep_nmax);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

double qr_split_err(const rbc_exact *h) {
  const double up = 0x1p-23, N = 3.0 * h->ldb + 1.0;
  return 3.1 * 0x1p-16 + 1.02 * N * up / (1.0 - N * up);
}

template <int KS>
static void launch_qr(const rbc_exact *h, const uint16_t *qhi, const uint16_t *qlo, uint32_t m, const float *qn2,
                      float *D, uint32_t grid, uint32_t nqt, uint32_t nrt, hipStream_t st) {
  hipLaunchKernelGGL((qr_mfma<KS>), dim3(grid), dim3(256), 0, st, qhi, qlo, m, h->reps_b.get(), h->reps_lo.get(),
                     h->ldb, h->reps_n2.get(), h->R, qn2, D, nqt, nrt);
}

int bf16_qr_dist(const rbc_exact *h, const uint16_t *qhi, const uint16_t *qlo, uint32_t m, const float *qn2, float *D,
                 hipStream_t st) {
  RBC_ARG_CHECK(h->metric == RBC_L2 && h->ldb <= 128 && h->reps_lo.get(), "split-BF16 q->rep distances need L2, d <= 128");
  const uint32_t nqt = cdiv(m, kQM), nrt = cdiv(h->R, kQM);
  const uint32_t grid = cdiv(nqt, kXCD) * kXCD * nrt;
  switch (h->ldb / 32) {
    case 1: launch_qr<1>(h, qhi, qlo, m, qn2, D, grid, nqt, nrt, st); break;
    case 2: launch_qr<2>(h, qhi, qlo, m, qn2, D, grid, nqt, nrt, st); break;
    case 3: launch_qr<3>(h, qhi, qlo, m, qn2, D, grid, nqt, nrt, st); break;
    default: launch_qr<4>(h, qhi, qlo, m, qn2, D, grid, nqt, nrt, st); break;
  }
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

// One persistent block per resident slot: (CUs x blocks per CU), a multiple of kXCD so that
// every XCD gets blocks.
template <int KS>
int launch_scan(const Bf16ScanArgs &a, uint32_t nitems, hipStream_t st) {
  static int per_cu = 0, cus = 0;
  if (per_cu == 0) {
    int dev = 0;
    RBC_HIP_TRY(hipGetDevice(&dev));
    RBC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RBC_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bf16_scan_kernel<KS>, 256, 0));
    per_cu = std::max(per_cu, 1);
  }
  uint32_t grid = (uint32_t)(cus * per_cu);
  grid = std::min(grid, nitems + kXCD);
  grid = cdiv(std::max(grid, (uint32_t)kXCD), kXCD) * kXCD;
  hipLaunchKernelGGL(bf16_scan_kernel<KS>, dim3(grid), dim3(256), 0, st, a);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int bf16_scan(rbc_exact *h, const uint16_t *qb, const float *qn, const float *qe, const double *gam,
              const uint32_t *qlist, const uint32_t *list_qoff, const ScanItem *items, uint32_t nitems,
              uint32_t *sched, float *cand_d, uint32_t *cand_i, uint32_t *cand_n, uint32_t cap,
              hipStream_t st) {
  if (nitems == 0) return RBC_OK;
  const uint32_t KS = h->ldb / 32;
  if (KS < 1 || KS > 4) {
    set_error("BF16 filter supports d <= 128 (d=%u)", h->d);
    return RBC_ERR_LIMIT;
  }
  hipLaunchKernelGGL((scan_split_kernel<kVQ, kVP, 16 * kQB>), dim3(1), dim3(1024), 0, st, list_qoff, h->list_len.get(), h->R, sched,
                     sched + kXCD + 1);
  RBC_HIP_TRY(hipGetLastError());
  Bf16ScanArgs a;
  a.items = items;
  a.bounds = sched;
  a.next = sched + kXCD + 1;
  a.qids = qlist;
  a.qb = qb;
  a.qn = qn;
  a.qe = qe;
  a.gam = gam;
  a.lqe = h->list_qerr.get();
  a.lnm = h->list_nmax.get();
  a.c_gemm = bf16_gemm_err(h);
  a.xb = h->xb.get();
  a.xn = h->xb_norm.get();
  a.perm = h->perm.get();
  a.n = h->n;
  a.ldb = h->ldb;
  a.cand_d = cand_d;
  a.cand_i = cand_i;
  a.cand_n = cand_n;
  a.cap = cap;
  switch (KS) {
    case 1: RBC_TRY(launch_scan<1>(a, nitems, st)); break;
    case 2: RBC_TRY(launch_scan<2>(a, nitems, st)); break;
    case 3: RBC_TRY(launch_scan<3>(a, nitems, st)); break;
    default: RBC_TRY(launch_scan<4>(a, nitems, st)); break;
  }
  return RBC_OK;
}

}  // namespace rbc

namespace rbc {

int bf16_probe(rbc_exact *h, const uint16_t *qb, const float *qn, const uint32_t *entries, uint32_t P,
               const ScanItem *items, uint32_t nitems, uint32_t Kp, uint32_t *out_i, float *out_v,
               hipStream_t st) {
  if (nitems == 0) return RBC_OK;
  if (Kp > (uint32_t)kPK) {
    set_error("BF16 probe keeps at most %d candidates (K'=%u)", kPK, Kp);
    return RBC_ERR_LIMIT;
  }
  Bf16ProbeArgs a;
  a.items = items;
  a.nitems = nitems;
  a.qids = entries;
  a.qb = qb;
  a.qn = qn;
  a.xb = h->xb.get();
  a.xn = h->xb_norm.get();
  a.perm = h->perm.get();
  a.ldb = h->ldb;
  a.Kp = Kp;
  a.P = P;
  // Slot bits: 8 slots per 32-point step of the longest list.
  uint32_t max_len = 0;
  for (uint32_t r = 0; r < h->R; ++r) max_len = std::max(max_len, h->h_list_off[r + 1] - h->h_list_off[r]);
  const uint32_t slots = std::max(1u, (max_len + kVP - 1) / kVP) * 8;
  uint32_t bits = 3;
  while ((1u << bits) < slots) ++bits;
  if (bits > 23) {  // lists of more than 2^25 points
    set_error("BF16 probe: a list of %u points is too long for its keys", max_len);
    return RBC_ERR_LIMIT;
  }
  a.idx_bits = bits;
  a.out_i = out_i;
  a.out_v = out_v;
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
