// Learned list proposal: a small MLP (d -> H1 -> H2 -> R, GELU, FP16 weights and
// activations, FP32 accumulation) trained on the GPU with the engine's own kernels.
//
// It predicts which ownership list holds a query's nearest neighbour.  The exact
// query uses its top-1 list as an extra first probe: a better probe tightens the
// bound on the k-th NN distance, so the plan prunes more lists and the scan emits
// fewer candidates.  It only changes which lists are probed first -- exactness is
// untouched (pruning and the re-rank do not depend on it).
//
// Kernels: a tiled FP16 GEMM on v_mfma_f32_16x16x32_f16 (64x64 tiles, 4 waves of 32x32,
// LDS-staged operands, generic strides so one kernel serves X W^T, dY^T X and dY W),
// with bias and GELU fused in its epilogue; softmax cross-entropy; GELU backward;
// column sums; Adam on FP32 master weights with FP16 copies.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "exact.hpp"

struct rbc_mlp;

namespace rbc {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

struct GemmArgs {
  const void *A;
  int64_t a_rs, a_cs;  // A(i,k) = A[i*a_rs + k*a_cs]
  const void *B;
  int64_t b_rs, b_cs;  // B(k,j) = B[k*b_rs + j*b_cs]
  float *C;            // FP32 result (pre-activation), C[i*ldc + j]; may be NULL
  int64_t ldc;
  _Float16 *H;         // FP16 copy of act(result), H[i*ldh + j]; may be NULL
  int64_t ldh;
  const float *bias;   // [N] or NULL
  int M, N, K;
  int gelu;            // apply GELU for H
  float alpha;         // result = alpha * A B + bias (0 means 1)
  unsigned long long *amax;  // if set: per-row argmax keys (argmax_key) instead of C and H
};

// Row argmax as one 64-bit maximum: ordered value bits above the complemented column, so
// the largest value wins and, among equal values, the lowest column; NaN (and no value at
// all) is key 0.
__device__ __forceinline__ unsigned long long argmax_key(float v, uint32_t j) {
  if (v != v) return 0ull;
  uint32_t u = __float_as_uint(v == 0.f ? 0.f : v);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)u << 32) | (0xFFFFFFFFu - j);
}

constexpr int kGT = 64, kGK = 32, kLdsPad = 40;

template <typename T>
__device__ __forceinline__ _Float16 to_h(T v) {
  return (_Float16)v;
}

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

// Result of wave w's 32 x 32 quarter of the 64 x 64 tile at (i0, j0): lane l holds rows
// rb + bi*16 + 4(l/16) + r, column cb + bj*16 + l%16 in acc[bi][bj][r].
__device__ __forceinline__ void gemm_epilogue(const GemmArgs &g, const f32x4 (&acc)[2][2], int i0, int j0, int w,
                                              int l) {
  const int rb = (w / 2) * 32, cb = (w % 2) * 32;
  if (g.amax) {
    // The 16 lanes l/16 = c hold rows rb + bi*16 + 4c + r; each reduces its two columns,
    // then the 16 reduce among themselves and one of them folds the row into amax.
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + rb + bi * 16 + (l / 16) * 4 + r;
        unsigned long long best = 0ull;
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const int j = j0 + cb + bj * 16 + l % 16;
          if (i < g.M && j < g.N) {
            const float v = (g.alpha != 0.f ? g.alpha : 1.f) * acc[bi][bj][r] + (g.bias ? g.bias[j] : 0.f);
            best = max(best, argmax_key(v, (uint32_t)j));
          }
        }
#pragma unroll
        for (int sh = 1; sh < 16; sh <<= 1) best = max(best, __shfl_xor(best, sh, 64));
        if (l % 16 == 0 && i < g.M && best) atomicMax(&g.amax[i], best);
      }
    return;
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int i = i0 + rb + bi * 16 + (l / 16) * 4 + r, j = j0 + cb + bj * 16 + l % 16;
        if (i >= g.M || j >= g.N) continue;
        float v = (g.alpha != 0.f ? g.alpha : 1.f) * acc[bi][bj][r] + (g.bias ? g.bias[j] : 0.f);
        if (g.C) g.C[i * g.ldc + j] = v;
        if (g.H) g.H[i * g.ldh + j] = (_Float16)(g.gelu ? gelu_f(v) : v);
      }
}

// KC: operands contiguous along k (a_cs == b_rs == 1, 16-byte aligned rows; the forward
// pass), B FP16, A FP16 or FP32 (the queries of the first layer, rounded to FP16 as they
// are staged, as the general kernel rounds them).  64-wide k chunks: every thread stages
// two 8-element pieces of one row of A and of B per chunk, and issues the next chunk's
// loads before the current chunk's MFMAs so that their latency overlaps the math.  Row
// pitch 72 halves: the 16 rows of a read pass fall on distinct banks.
constexpr int kKC = 64, kKCPad = kKC + 8;

template <typename TA>
__device__ __forceinline__ f16x8 load8_h(const TA *p) {
  if constexpr (sizeof(TA) == 2) {
    return *reinterpret_cast<const f16x8 *>(p);
  } else {
    const float4 u = reinterpret_cast<const float4 *>(p)[0], v = reinterpret_cast<const float4 *>(p)[1];
    return f16x8{(_Float16)u.x, (_Float16)u.y, (_Float16)u.z, (_Float16)u.w,
                 (_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
  }
}

template <typename TA>
__global__ __launch_bounds__(256) void gemm_f16_kc(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) _Float16 As[kGT][kKCPad];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[kGT][kKCPad];  // Bs[j][k]
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int i0 = blockIdx.y * kGT, j0 = blockIdx.x * kGT;
  const TA *A = static_cast<const TA *>(g.A);
  const _Float16 *B = static_cast<const _Float16 *>(g.B);
  const int lr = t / 4, lk = (t % 4) * 8;  // row; k offsets lk and lk + 32 of the chunk
  const int i = i0 + lr, j = j0 + lr;
  f16x8 ra[2], rb[2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int k = k0 + lk + 32 * e;
      ra[e] = (i < g.M && k < g.K) ? load8_h(A + int64_t(i) * g.a_rs + k) : f16x8{};
      rb[e] = (j < g.N && k < g.K) ? *reinterpret_cast<const f16x8 *>(B + int64_t(j) * g.b_cs + k) : f16x8{};
    }
  };
  f32x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int rb0 = (w / 2) * 32, cb0 = (w % 2) * 32;
  gload(0);
  for (int k0 = 0; k0 < g.K; k0 += kKC) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      *reinterpret_cast<f16x8 *>(&As[lr][lk + 32 * e]) = ra[e];
      *reinterpret_cast<f16x8 *>(&Bs[lr][lk + 32 * e]) = rb[e];
    }
    __syncthreads();
    if (k0 + kKC < g.K) gload(k0 + kKC);
#pragma unroll
    for (int ks = 0; ks < kKC / 32; ++ks)
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) {
        const f16x8 av = *reinterpret_cast<const f16x8 *>(&As[rb0 + bi * 16 + l % 16][ks * 32 + (l / 16) * 8]);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const f16x8 bv = *reinterpret_cast<const f16x8 *>(&Bs[cb0 + bj * 16 + l % 16][ks * 32 + (l / 16) * 8]);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, acc[bi][bj], 0, 0, 0);
        }
      }
  }
  gemm_epilogue(g, acc, i0, j0, w, l);
}

// Last layer with the row argmax fused (inference): only argmax_j (alpha A W^T + b)(i, j)
// is kept, as an argmax_key per row.  A block owns 64 rows and walks all N columns in
// 64-column tiles: wave w keeps rows (w/2)*32.. of A in registers (2 x KS MFMA fragments),
// the tile of W (64 rows of W, K = 32 KS) is staged in LDS once per block and shared by the
// 4 waves, the next tile's loads are issued before the current tile's MFMAs, and every
// lane keeps the best key of each of its 8 rows.  Row pitch CPR + 1 chunks (CPR even): the
// 16 rows of a read pass fall on distinct banks.
template <int KS>
__global__ __launch_bounds__(256) void gemm_f16_argmax(GemmArgs g) {
  constexpr int K = 32 * KS, CPR = K / 8, PITCH = CPR + 1;
  __shared__ f16x8 Ws[kGT * PITCH];
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int i0 = blockIdx.x * kGT, rb = (w / 2) * 32, cb = (w % 2) * 32;
  const _Float16 *A = static_cast<const _Float16 *>(g.A);
  const _Float16 *W = static_cast<const _Float16 *>(g.B);  // column j of B = row j of W
  const float alpha = g.alpha != 0.f ? g.alpha : 1.f;
  f16x8 af[2][KS];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
    const int i = i0 + rb + bi * 16 + l % 16;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      af[bi][ks] = i < g.M ? *reinterpret_cast<const f16x8 *>(A + int64_t(i) * g.a_rs + ks * 32 + (l / 16) * 8)
                           : f16x8{};
  }
  unsigned long long best[2][4];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) best[bi][r] = 0ull;
  // Staging: chunk c = t + 256 e of the tile is chunk c % CPR of W row j0 + c / CPR.
  f16x8 wr[KS];
  auto wload = [&](int j0) {
#pragma unroll
    for (int e = 0; e < KS; ++e) {
      const int c = t + 256 * e, j = j0 + c / CPR;
      wr[e] = j < g.N ? *reinterpret_cast<const f16x8 *>(W + int64_t(j) * g.b_cs + (c % CPR) * 8) : f16x8{};
    }
  };
  wload(0);
  for (int j0 = 0; j0 < g.N; j0 += kGT) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < KS; ++e) {
      const int c = t + 256 * e;
      Ws[(c / CPR) * PITCH + c % CPR] = wr[e];
    }
    __syncthreads();
    if (j0 + kGT < g.N) wload(j0 + kGT);
    f32x4 acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const f16x8 bv = Ws[(cb + bj * 16 + l % 16) * PITCH + ks * 4 + l / 16];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[bi][ks], bv, acc[bi][bj], 0, 0, 0);
      }
    // Lane l holds rows rb + bi*16 + 4(l/16) + r of columns j0 + cb + bj*16 + l%16.
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int j = j0 + cb + bj * 16 + l % 16;
      if (j < g.N) {
        const float bias = g.bias ? g.bias[j] : 0.f;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int r = 0; r < 4; ++r) best[bi][r] = max(best[bi][r], argmax_key(alpha * acc[bi][bj][r] + bias, j));
      }
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      unsigned long long b = best[bi][r];
#pragma unroll
      for (int sh = 1; sh < 16; sh <<= 1) b = max(b, __shfl_xor(b, sh, 64));
      const int i = i0 + rb + bi * 16 + (l / 16) * 4 + r;
      if (l % 16 == 0 && i < g.M && b) atomicMax(&g.amax[i], b);
    }
}

// The smallest float above x (x not NaN): -FLT_MAX for -inf, the least denormal for +-0.
__device__ __forceinline__ float next_float_up(float x) {
  if (x == -INFINITY) return -FLT_MAX;
  if (x == 0.f) return __int_as_float(1);
  const int b = __float_as_int(x);
  return __int_as_float(b >= 0 ? b + 1 : b - 1);
}

// Last layer with the top-P lists fused (inference): for every row i, the Pr <= P largest
// values of (A W^T + b)(i, :) (descending, lowest column among equal values; NaN counts as
// -inf; alpha must be 0 or 1) go to lists[i*Pr ..], and count[i] gets how many of them it
// takes for their softmax probabilities to reach tau (0 if all Pr fall short), so the
// logits never leave the chip.  A block owns kToppRows = 128 rows (wave w: rows 32 w ..
// +31, as two 16-row MFMA blocks whose fragments stay in registers) and walks all N
// columns in 64-column tiles; the tile of W and its biases are staged in LDS once for the
// 128 rows.  The MFMA operands are swapped, so lane l's accumulators hold 4 columns of ONE
// row (row l%16 of a row block): the 4 lanes l%16, +16, +32, +48 of a wave together see
// every column of a row, and each keeps its P best (value, column) sorted in registers and
// an online softmax normaliser (running max, sum of exponentials).  The accumulators start
// at the columns' biases (-inf past the last column, so no bounds test per value).  The 4
// partial lists of a row are merged by one thread per row through LDS.
constexpr int kToppRows = 128;

template <int KS, int P>
__global__ __launch_bounds__(256) void gemm_f16_topp(GemmArgs g, int Pr, float tau, uint32_t *lists,
                                                     uint32_t *count) {
  constexpr int K = 32 * KS, CPR = K / 8, PITCH = CPR + 1;
  constexpr int kWsBytes = kGT * PITCH * 16;
  constexpr int kMergeBytes = kToppRows * 4 * P * 8 + kToppRows * 4 * 8;
  // The W tile during the column walk; the partial lists and normalisers after it.
  __shared__ __attribute__((aligned(16))) unsigned char smem[kWsBytes > kMergeBytes ? kWsBytes : kMergeBytes];
  __shared__ __attribute__((aligned(16))) float bs[kGT];
  f16x8 *Ws = reinterpret_cast<f16x8 *>(smem);
  float(*mv)[4][P] = reinterpret_cast<float(*)[4][P]>(smem);
  int(*mi)[4][P] = reinterpret_cast<int(*)[4][P]>(smem + kToppRows * 4 * P * 4);
  float(*pm)[4] = reinterpret_cast<float(*)[4]>(smem + kToppRows * 4 * P * 8);
  float(*ps)[4] = reinterpret_cast<float(*)[4]>(smem + kToppRows * 4 * P * 8 + kToppRows * 4 * 4);
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int i0 = blockIdx.x * kToppRows, rb = w * 32;
  const _Float16 *A = static_cast<const _Float16 *>(g.A);
  const _Float16 *W = static_cast<const _Float16 *>(g.B);  // column j of B = row j of W
  f16x8 af[2][KS];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
    const int i = i0 + rb + bi * 16 + l % 16;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      af[bi][ks] = i < g.M ? *reinterpret_cast<const f16x8 *>(A + int64_t(i) * g.a_rs + ks * 32 + (l / 16) * 8)
                           : f16x8{};
  }
  float tv[2][P], mx[2], se[2];
  int ti[2][P];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
    mx[bi] = -INFINITY;
    se[bi] = 0.f;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      tv[bi][s] = -INFINITY;
      ti[bi][s] = 0x7FFFFFFF;
    }
  }
  // Tiles of W are fetched two tiles ahead into two register sets that alternate (the tile
  // loop is written out twice per iteration, so each set keeps fixed registers).
  f16x8 wr0[KS], wr1[KS];
  float br0 = 0.f, br1 = 0.f;
  auto wload = [&](int j0, f16x8 (&wr)[KS], float &br) {
#pragma unroll
    for (int e = 0; e < KS; ++e) {
      const int c = t + 256 * e, j = j0 + c / CPR;
      wr[e] = j < g.N ? *reinterpret_cast<const f16x8 *>(W + int64_t(j) * g.b_cs + (c % CPR) * 8) : f16x8{};
    }
    if (t < kGT) br = j0 + t < g.N ? (g.bias ? g.bias[j0 + t] : 0.f) : -INFINITY;
  };
  auto tile = [&](int j0, f16x8 (&wr)[KS], float &br) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < KS; ++e) {
      const int c = t + 256 * e;
      Ws[(c / CPR) * PITCH + c % CPR] = wr[e];
    }
    if (t < kGT) bs[t] = br;
    __syncthreads();
    if (j0 + 2 * kGT < g.N) wload(j0 + 2 * kGT, wr, br);
    f32x4 acc[2][4];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) {
      const f32x4 b4 = *reinterpret_cast<const f32x4 *>(&bs[bj * 16 + (l / 16) * 4]);
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) acc[bi][bj] = b4;
    }
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        const f16x8 wv = Ws[(bj * 16 + l % 16) * PITCH + ks * 4 + l / 16];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wv, af[bi][ks], acc[bi][bj], 0, 0, 0);
      }
    // Lane l holds row rb + bi*16 + l%16, columns j0 + bj*16 + 4(l/16) + r: value e = 4 bj
    // + r of a row block is column j0 + 16 (e/4) + 4(l/16) + e%4, so a lane's values come
    // in ascending column order.
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
      float v[16];
      float m16 = -INFINITY;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        v[e] = fmaxf(acc[bi][e / 4][e % 4], -INFINITY);  // v_max_f32: a NaN becomes -inf
        m16 = fmaxf(m16, v[e]);
      }
      if (m16 > mx[bi]) {
        se[bi] = mx[bi] == -INFINITY ? 0.f : se[bi] * __expf(mx[bi] - m16);
        mx[bi] = m16;
      }
      if (mx[bi] != -INFINITY) {
        // exp(v - mx) as 2^(v log2 e - mx log2 e), one FMA per value; exp2(-inf) = 0.
        const float ml = mx[bi] * 1.4426950408889634f;
#pragma unroll
        for (int e = 0; e < 16; ++e)
          se[bi] += __builtin_amdgcn_exp2f(__builtin_fmaf(v[e], 1.4426950408889634f, -ml));
      }
      // The 4 lanes of a row (l, l^16, l^32, l^48) each hold P values >= their own P-th,
      // so a value below another lane's P-th is not among the row's P best; nor is one
      // below the smallest of the 4 lanes' (P/4)-th values (those lanes hold P distinct
      // columns at or above it).  A value equal to such a bound may be (a lower column
      // wins a tie); one equal to the lane's own P-th is not (the lane's entries have
      // lower columns).  Only values that pass these tests are inserted, one per lane per
      // pass, so a wave runs as many insertion passes as its busiest lane has hits rather
      // than one per value.
      const float own = tv[bi][P - 1], q4 = tv[bi][P / 4 - 1];
      const float oth = fmaxf(
          fmaxf(__shfl_xor(own, 16, 64), fmaxf(__shfl_xor(own, 32, 64), __shfl_xor(own, 48, 64))),
          fminf(fminf(q4, __shfl_xor(q4, 16, 64)), fminf(__shfl_xor(q4, 32, 64), __shfl_xor(q4, 48, 64))));
      // v > own and v >= oth in one compare: v >= max(the next float above own, oth).
      const float thr = fmaxf(next_float_up(own), oth);
      uint32_t pend = 0;
#pragma unroll
      for (int e = 0; e < 16; ++e) pend |= v[e] >= thr ? (1u << e) : 0u;
      while (pend) {
        const int e = __builtin_ctz(pend);
        pend &= pend - 1;
        float x = v[0];
#pragma unroll
        for (int s = 1; s < 16; ++s) x = e == s ? v[s] : x;
        if (!(x > tv[bi][P - 1])) continue;
        const int j = j0 + (e / 4) * 16 + (l / 16) * 4 + e % 4;
        // Branch-free insertion from the bottom: slot s takes slot s-1's entry when x
        // belongs above it, x itself when it belongs exactly there, else keeps its own.
#pragma unroll
        for (int s = P - 1; s > 0; --s) {
          const bool above = x > tv[bi][s - 1];
          const bool here = !above && x > tv[bi][s];
          const float nv = above ? tv[bi][s - 1] : (here ? x : tv[bi][s]);
          const int ni = above ? ti[bi][s - 1] : (here ? j : ti[bi][s]);
          tv[bi][s] = nv;
          ti[bi][s] = ni;
        }
        if (x > tv[bi][0]) {
          tv[bi][0] = x;
          ti[bi][0] = j;
        }
      }
    }
  };
  wload(0, wr0, br0);
  if (kGT < g.N) wload(kGT, wr1, br1);
  for (int j0 = 0; j0 < g.N; j0 += 2 * kGT) {
    tile(j0, wr0, br0);
    if (j0 + kGT < g.N) tile(j0 + kGT, wr1, br1);
  }
  // Partial p = l/16 of row rb + bi*16 + l%16 (the W tile's LDS is reused).
  __syncthreads();
  const int part = l / 16;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
    const int rr = rb + bi * 16 + l % 16;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      mv[rr][part][s] = tv[bi][s];
      mi[rr][part][s] = ti[bi][s];
    }
    pm[rr][part] = mx[bi];
    ps[rr][part] = se[bi];
  }
  __syncthreads();
  if (t >= kToppRows || i0 + t >= g.M) return;
  const int rr = t, i = i0 + t;
  float M = -INFINITY;
  for (int q = 0; q < 4; ++q) M = fmaxf(M, pm[rr][q]);
  float S = 0.f;
  for (int q = 0; q < 4; ++q)
    if (pm[rr][q] != -INFINITY) S += ps[rr][q] * __expf(pm[rr][q] - M);
  int head[4] = {0, 0, 0, 0};
  float mass = 0.f;
  uint32_t cnt = 0;
  for (int r = 0; r < Pr; ++r) {
    int bq = -1;
    float bv = -INFINITY;
    int bidx = 0x7FFFFFFF;
    for (int q = 0; q < 4; ++q) {
      if (head[q] >= P) continue;
      const float v = mv[rr][q][head[q]];
      const int j = mi[rr][q][head[q]];
      if (j == 0x7FFFFFFF) continue;
      if (bq < 0 || v > bv || (v == bv && j < bidx)) {
        bq = q;
        bv = v;
        bidx = j;
      }
    }
    if (bq >= 0) ++head[bq];
    lists[int64_t(i) * Pr + r] = bq >= 0 ? (uint32_t)bidx : 0xFFFFFFFFu;
    if (cnt == 0 && bq >= 0 && S > 0.f) {
      mass += __expf(bv - M) / S;
      if (mass >= tau) cnt = r + 1;
    }
  }
  count[i] = cnt;
}

// General operands (any element strides, FP32 or FP16): 32-wide k chunks staged element by
// element.
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void gemm_f16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) _Float16 As[kGT][kLdsPad];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[kGT][kLdsPad];  // Bs[j][k]
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int i0 = blockIdx.y * kGT, j0 = blockIdx.x * kGT;
  const TA *A = static_cast<const TA *>(g.A);
  const TB *B = static_cast<const TB *>(g.B);
  f32x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lr = t / 4, lk = (t % 4) * 8;
  for (int k0 = 0; k0 < g.K; k0 += kGK) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int i = i0 + lr, k = k0 + lk + e, j = j0 + lr;
      As[lr][lk + e] = (i < g.M && k < g.K) ? to_h(A[i * g.a_rs + k * g.a_cs]) : (_Float16)0.f;
      Bs[lr][lk + e] = (j < g.N && k < g.K) ? to_h(B[k * g.b_rs + j * g.b_cs]) : (_Float16)0.f;
    }
    __syncthreads();
    const int rb = (w / 2) * 32, cb = (w % 2) * 32;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
      f16x8 av = *reinterpret_cast<const f16x8 *>(&As[rb + bi * 16 + l % 16][(l / 16) * 8]);
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        f16x8 bv = *reinterpret_cast<const f16x8 *>(&Bs[cb + bj * 16 + l % 16][(l / 16) * 8]);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, acc[bi][bj], 0, 0, 0);
      }
    }
  }
  gemm_epilogue(g, acc, i0, j0, w, l);
}

template <typename TA, typename TB>
int gemm(const GemmArgs &g, hipStream_t st) {
  dim3 grid(cdiv(g.N, kGT), cdiv(g.M, kGT));
  const bool kc = sizeof(TB) == 2 && g.a_cs == 1 && g.b_rs == 1 && g.K % 8 == 0 &&
                  (g.a_rs * (int64_t)sizeof(TA)) % 16 == 0 && g.b_cs % 8 == 0 &&
                  reinterpret_cast<uintptr_t>(g.A) % 16 == 0 && reinterpret_cast<uintptr_t>(g.B) % 16 == 0;
  if (kc && sizeof(TA) == 2 && g.amax && g.K % 32 == 0 && g.K <= 256) {
    const dim3 rows(cdiv(g.M, kGT));
    switch (g.K / 32) {
      case 1: hipLaunchKernelGGL(gemm_f16_argmax<1>, rows, dim3(256), 0, st, g); break;
      case 2: hipLaunchKernelGGL(gemm_f16_argmax<2>, rows, dim3(256), 0, st, g); break;
      case 3: hipLaunchKernelGGL(gemm_f16_argmax<3>, rows, dim3(256), 0, st, g); break;
      case 4: hipLaunchKernelGGL(gemm_f16_argmax<4>, rows, dim3(256), 0, st, g); break;
      case 5: hipLaunchKernelGGL(gemm_f16_argmax<5>, rows, dim3(256), 0, st, g); break;
      case 6: hipLaunchKernelGGL(gemm_f16_argmax<6>, rows, dim3(256), 0, st, g); break;
      case 7: hipLaunchKernelGGL(gemm_f16_argmax<7>, rows, dim3(256), 0, st, g); break;
      default: hipLaunchKernelGGL(gemm_f16_argmax<8>, rows, dim3(256), 0, st, g); break;
    }
  } else if (kc)
    hipLaunchKernelGGL(gemm_f16_kc<TA>, grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((gemm_f16<TA, TB>), grid, dim3(256), 0, st, g);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

// Softmax cross-entropy of rows of logits [B][N] against soft targets: row i's target is
// uniform over its valid labels labels[i*L .. i*L+L) (a label >= N is no label; with L = 1
// this is the one-hot cross-entropy).  dZ = softmax - target as FP16 (the 1/B of the mean
// is applied to the weight gradients in FP32, so that FP16 does not have to hold values of
// order 1/(B N)); the mean loss sum_l (logsumexp - z_label_l) / L_valid over the B rows
// accumulates into *loss.  A row without a valid label contributes neither.  One wave per
// row; the target of wave w is built in LDS (N floats per wave).
__global__ __launch_bounds__(256) void softmax_ce(const float *Z, const uint32_t *labels, int L, int B, int N,
                                                  _Float16 *dZ, float *loss) {
  extern __shared__ float tgt_all[];
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int row = blockIdx.x * (blockDim.x / 64) + w;
  if (row >= B) return;
  float *tgt = tgt_all + size_t(w) * N;
  const float *z = Z + int64_t(row) * N;
  for (int j = l; j < N; j += 64) tgt[j] = 0.f;
  uint32_t lab = l < L ? labels[int64_t(row) * L + l] : 0xFFFFFFFFu;
  const bool valid = lab < (uint32_t)N;
  const int nv = __popcll(__ballot(valid));
  float mx = -INFINITY;
  for (int j = l; j < N; j += 64) mx = fmaxf(mx, z[j]);
  for (int s = 32; s > 0; s >>= 1) mx = fmaxf(mx, __shfl_xor(mx, s, 64));
  float se = 0.f;
  for (int j = l; j < N; j += 64) se += __expf(z[j] - mx);
  for (int s = 32; s > 0; s >>= 1) se += __shfl_xor(se, s, 64);
  const float share = nv ? 1.f / (float)nv : 0.f;
  __builtin_amdgcn_wave_barrier();
  if (valid) atomicAdd(&tgt[lab], share);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  for (int j = l; j < N; j += 64) {
    const float p = __expf(z[j] - mx) / se;
    dZ[int64_t(row) * N + j] = (_Float16)(nv ? p - tgt[j] : 0.f);
  }
  float lv = valid ? (logf(se) + mx - z[lab]) * share : 0.f;
  for (int s = 32; s > 0; s >>= 1) lv += __shfl_xor(lv, s, 64);
  if (l == 0 && loss && nv) atomicAdd(loss, lv / (float)B);
}

// dZ = dH * gelu'(Z), FP16 out.
__global__ void gelu_backward(const float *dH, const float *Z, int64_t n, _Float16 *dZ) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dZ[i] = (_Float16)(dH[i] * gelu_grad(Z[i]));
}

// db[j] = scale * sum_i dZ[i][j].  A block of 1024 threads owns 64 columns: 16 row
// groups each sum every 16th row, then the 16 partials are added in a fixed order.
__global__ __launch_bounds__(1024) void col_sums(const _Float16 *dZ, int B, int N, float scale, float *db) {
  __shared__ float part[16][64];
  const int c = threadIdx.x % 64, g = threadIdx.x / 64;
  const int j = blockIdx.x * 64 + c;
  float s = 0.f;
  if (j < N) {
#pragma unroll 8
    for (int i = g; i < B; i += 16) s += (float)dZ[int64_t(i) * N + j];
  }
  part[g][c] = s;
  __syncthreads();
  if (g == 0 && j < N) {
    float t = 0.f;
    for (int r = 0; r < 16; ++r) t += part[r][c];
    db[j] = scale * t;
  }
}

__global__ void adam_step(float *w, _Float16 *w16, const float *g, float *m1, float *m2, int64_t n, float lr,
                          float b1, float b2, float eps, float c1, float c2) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float a = b1 * m1[i] + (1.f - b1) * gi;
  float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  float v = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[i] = v;
  if (w16) w16[i] = (_Float16)v;
}

__global__ void to_half(const float *x, int64_t n, _Float16 *y) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (_Float16)x[i];
}

// Gathers rows idx[0..B) of X (FP32, ldx) into an FP16 [B][d] batch.
// FP16 copy of rows idx[0..B) of X (rows 0..B when idx is null).
__global__ void gather_half(const float *X, int64_t ldx, const uint32_t *idx, int B, int d, _Float16 *out) {
  int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(B) * d) return;
  int r = int(e / d), c = int(e % d);
  out[e] = (_Float16)X[int64_t(idx ? idx[r] : uint32_t(r)) * ldx + c];
}

__global__ void argmax_decode(const unsigned long long *key, uint32_t B, uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B) out[i] = key[i] ? 0xFFFFFFFFu - (uint32_t)key[i] : 0xFFFFFFFFu;
}

// out[i][0..L) = src[idx[i]][0..L).
__global__ void gather_u32(const uint32_t *src, const uint32_t *idx, int B, int L, uint32_t *out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < B * L) out[e] = src[int64_t(idx[e / L]) * L + e % L];
}

}  // namespace

}  // namespace rbc

// Parameters: W1 [H1][D], b1 [H1], W2 [H2][H1], b2 [H2], W3 [C][H2], b3 [C] in one FP32
// buffer (this order) with an FP16 copy.
struct rbc_mlp {
  uint32_t D = 0, H1 = 0, H2 = 0, C = 0;
  size_t np = 0;
  size_t off[7];
  rbc::DevBuf<float> w, g, m1, m2;
  rbc::DevBuf<_Float16> w16;
  uint64_t step = 0;
  // Activations for a batch of up to `cap` rows.
  uint32_t cap = 0;
  rbc::DevBuf<_Float16> x16, h1, h2, dz3, dz2, dz1;
  rbc::DevBuf<float> z1, z2, z3, dh2, dh1, loss;
  rbc::DevBuf<uint32_t> lab, idx;
  rbc::DevBuf<unsigned long long> amax;  // inference: per-row argmax keys
  uint32_t labcap = 0;                   // label entries m->lab holds (rows x labels per row)
  // Inference activations (FP16 hidden layers only), for up to icap rows.
  uint32_t icap = 0;
  rbc::DevBuf<_Float16> ih1, ih2;

  int ensure_infer(uint32_t B) {
    if (B <= icap) return RBC_OK;
    RBC_TRY(ih1.alloc(size_t(B) * H1));
    RBC_TRY(ih2.alloc(size_t(B) * H2));
    icap = B;
    return RBC_OK;
  }

  int ensure_labels(size_t n) {
    if (n <= labcap) return RBC_OK;
    RBC_TRY(lab.alloc(n));
    labcap = (uint32_t)n;
    return RBC_OK;
  }

  int ensure(uint32_t B) {
    if (B <= cap) return RBC_OK;
    cap = B;
    RBC_TRY(x16.alloc(size_t(B) * D));
    RBC_TRY(h1.alloc(size_t(B) * H1));
    RBC_TRY(h2.alloc(size_t(B) * H2));
    RBC_TRY(dz3.alloc(size_t(B) * C));
    RBC_TRY(dz2.alloc(size_t(B) * H2));
    RBC_TRY(dz1.alloc(size_t(B) * H1));
    RBC_TRY(z1.alloc(size_t(B) * H1));
    RBC_TRY(z2.alloc(size_t(B) * H2));
    RBC_TRY(z3.alloc(size_t(B) * C));
    RBC_TRY(dh2.alloc(size_t(B) * H2));
    RBC_TRY(dh1.alloc(size_t(B) * H1));
    RBC_TRY(ensure_labels(B));
    RBC_TRY(idx.alloc(B));
    RBC_TRY(amax.alloc(B));
    RBC_TRY(loss.alloc(1));
    return RBC_OK;
  }
  float *W(int i) { return w.get() + off[2 * i]; }
  float *Bv(int i) { return w.get() + off[2 * i + 1]; }
  _Float16 *W16(int i) { return w16.get() + off[2 * i]; }
  float *GW(int i) { return g.get() + off[2 * i]; }
  float *GB(int i) { return g.get() + off[2 * i + 1]; }
};

namespace rbc {

namespace {

// Forward pass on the FP16 batch in m->x16 (B rows); logits to m->z3 (or `logits`).
int mlp_forward(rbc_mlp *m, uint32_t B, float *logits, hipStream_t st, unsigned long long *amax = nullptr) {
  GemmArgs g{};
  g.A = m->x16.get(); g.a_rs = m->D; g.a_cs = 1;
  g.B = m->W16(0); g.b_rs = 1; g.b_cs = m->D;
  g.C = m->z1.get(); g.ldc = m->H1; g.H = m->h1.get(); g.ldh = m->H1; g.bias = m->Bv(0);
  g.M = B; g.N = m->H1; g.K = m->D; g.gelu = 1;
  RBC_TRY((gemm<_Float16, _Float16>(g, st)));
  g = GemmArgs{};
  g.A = m->h1.get(); g.a_rs = m->H1; g.a_cs = 1;
  g.B = m->W16(1); g.b_rs = 1; g.b_cs = m->H1;
  g.C = m->z2.get(); g.ldc = m->H2; g.H = m->h2.get(); g.ldh = m->H2; g.bias = m->Bv(1);
  g.M = B; g.N = m->H2; g.K = m->H1; g.gelu = 1;
  RBC_TRY((gemm<_Float16, _Float16>(g, st)));
  g = GemmArgs{};
  g.A = m->h2.get(); g.a_rs = m->H2; g.a_cs = 1;
  g.B = m->W16(2); g.b_rs = 1; g.b_cs = m->H2;
  g.C = amax ? nullptr : logits ? logits : m->z3.get(); g.ldc = m->C; g.bias = m->Bv(2);
  g.M = B; g.N = m->C; g.K = m->H2; g.gelu = 0; g.amax = amax;
  RBC_TRY((gemm<_Float16, _Float16>(g, st)));
  return RBC_OK;
}

// Soft-target cross-entropy and its logit gradient (L labels per row in m->lab): waves per
// block as the LDS for the per-wave targets (N floats each) allows.
int launch_softmax_ce(rbc_mlp *m, uint32_t B, uint32_t L, hipStream_t st) {
  const size_t per_wave = size_t(m->C) * sizeof(float);
  const uint32_t waves = (uint32_t)std::min<size_t>(4, std::max<size_t>(1, 65536 / per_wave));
  if (waves * per_wave > 65536) {
    set_error("training supports at most 16384 outputs (n_out=%u)", m->C);
    return RBC_ERR_LIMIT;
  }
  hipLaunchKernelGGL(softmax_ce, dim3(cdiv(B, waves)), dim3(64 * waves), waves * per_wave, st, m->z3.get(),
                     m->lab.get(), (int)L, (int)B, (int)m->C, m->dz3.get(), m->loss.get());
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

// Backward pass for the batch (L labels per row in m->lab); gradients to m->g.
int mlp_backward(rbc_mlp *m, uint32_t B, uint32_t L, hipStream_t st) {
  RBC_TRY(fill_f32(m->loss.get(), 0.f, 1, st));
  RBC_TRY(launch_softmax_ce(m, B, L, st));
  const _Float16 *dz[3] = {m->dz1.get(), m->dz2.get(), m->dz3.get()};
  const _Float16 *in[3] = {m->x16.get(), m->h1.get(), m->h2.get()};
  const uint32_t outd[3] = {m->H1, m->H2, m->C}, ind[3] = {m->D, m->H1, m->H2};
  float *dh[3] = {nullptr, m->dh1.get(), m->dh2.get()};
  const float *zs[3] = {nullptr, m->z1.get(), m->z2.get()};
  _Float16 *dzo[3] = {nullptr, m->dz1.get(), m->dz2.get()};
  for (int layer = 2; layer >= 0; --layer) {
    // dW = dZ^T X  ([out][in]), db = column sums of dZ.
    GemmArgs g{};
    g.A = dz[layer]; g.a_rs = 1; g.a_cs = outd[layer];
    g.B = in[layer]; g.b_rs = ind[layer]; g.b_cs = 1;
    g.C = m->GW(layer); g.ldc = ind[layer];
    g.M = outd[layer]; g.N = ind[layer]; g.K = B;
    g.alpha = 1.f / (float)B;
    RBC_TRY((gemm<_Float16, _Float16>(g, st)));
    hipLaunchKernelGGL(col_sums, dim3(cdiv(outd[layer], 64)), dim3(1024), 0, st, dz[layer], (int)B,
                       (int)outd[layer], 1.f / (float)B, m->GB(layer));
    RBC_HIP_TRY(hipGetLastError());
    if (layer == 0) break;
    // dH_{layer} = dZ W ([B][in]) (per-row gradients, without the 1/B), then
    // dZ_{layer-1} = dH * gelu'(Z_{layer}).
    g = GemmArgs{};
    g.A = dz[layer]; g.a_rs = outd[layer]; g.a_cs = 1;
    g.B = m->W16(layer); g.b_rs = ind[layer]; g.b_cs = 1;
    g.C = dh[layer]; g.ldc = ind[layer];
    g.M = B; g.N = ind[layer]; g.K = outd[layer];
    RBC_TRY((gemm<_Float16, _Float16>(g, st)));
    int64_t n = int64_t(B) * ind[layer];
    hipLaunchKernelGGL(gelu_backward, dim3((uint32_t)cdivz(n, 256)), dim3(256), 0, st, dh[layer], zs[layer],
                       n, dzo[layer]);
    RBC_HIP_TRY(hipGetLastError());
  }
  return RBC_OK;
}

}  // namespace

// Logits -> top-1 list per row.
// Stream-ordered: out is complete when the work queued on st is.
int mlp_propose(rbc_mlp *m, const float *q, size_t ldq, uint32_t rows, uint32_t *out, hipStream_t st) {
  const uint32_t chunk = 65536;
  RBC_TRY(m->ensure(std::min(rows, chunk)));
  for (uint32_t r0 = 0; r0 < rows; r0 += chunk) {
    uint32_t B = std::min(chunk, rows - r0);
    int64_t ne = int64_t(B) * m->D;
    hipLaunchKernelGGL(gather_half, dim3((uint32_t)cdivz(ne, 256)), dim3(256), 0, st, q + size_t(r0) * ldq,
                       (int64_t)ldq, (const uint32_t *)nullptr, (int)B, (int)m->D, m->x16.get());
    RBC_HIP_TRY(hipGetLastError());
    RBC_TRY(fill_u32(reinterpret_cast<uint32_t *>(m->amax.get()), 0u, 2 * size_t(B), st));
    RBC_TRY(mlp_forward(m, B, nullptr, st, m->amax.get()));
    hipLaunchKernelGGL(argmax_decode, dim3(cdiv(B, 256)), dim3(256), 0, st, m->amax.get(), B, out + r0);
    RBC_HIP_TRY(hipGetLastError());
  }
  return RBC_OK;
}

// The P best lists of each row by logit and how many of them reach softmax mass tau
// (0 = not within P), into lists[rows][P] and count[rows].  Stream-ordered.  Inference
// keeps only what it needs: layer 1 reads the FP32 queries directly, the hidden layers
// store only their FP16 activations, and the last layer's logits stay on the chip (the
// top-P and the softmax mass are fused into its GEMM).
int mlp_propose_lists(rbc_mlp *m, const float *q, size_t ldq, uint32_t rows, uint32_t P, float tau,
                      uint32_t *lists, uint32_t *count, hipStream_t st) {
  if (P < 1 || P > 16) {
    set_error("a list proposal holds 1 to 16 lists (P=%u)", P);
    return RBC_ERR_ARG;
  }
  if (m->H2 % 32 != 0 || m->H2 > 256 || m->H1 % 8 != 0) {
    set_error("list proposals need hidden widths with H1 %% 8 == 0 and H2 %% 32 == 0, H2 <= 256 (%u, %u)",
              m->H1, m->H2);
    return RBC_ERR_LIMIT;
  }
  // One pass over up to 2^20 rows at a time (FP16 activations only: 1 KB a row at 256 + 256).
  const uint32_t chunk = 1u << 20;
  RBC_TRY(m->ensure_infer(std::min(rows, chunk)));
  for (uint32_t r0 = 0; r0 < rows; r0 += chunk) {
    const uint32_t B = std::min(chunk, rows - r0);
    GemmArgs g{};
    g.A = q + size_t(r0) * ldq; g.a_rs = (int64_t)ldq; g.a_cs = 1;
    g.B = m->W16(0); g.b_rs = 1; g.b_cs = m->D;
    g.H = m->ih1.get(); g.ldh = m->H1; g.bias = m->Bv(0);
    g.M = B; g.N = m->H1; g.K = m->D; g.gelu = 1;
    RBC_TRY((gemm<float, _Float16>(g, st)));
    g = GemmArgs{};
    g.A = m->ih1.get(); g.a_rs = m->H1; g.a_cs = 1;
    g.B = m->W16(1); g.b_rs = 1; g.b_cs = m->H1;
    g.H = m->ih2.get(); g.ldh = m->H2; g.bias = m->Bv(1);
    g.M = B; g.N = m->H2; g.K = m->H1; g.gelu = 1;
    RBC_TRY((gemm<_Float16, _Float16>(g, st)));
    g = GemmArgs{};
    g.A = m->ih2.get(); g.a_rs = m->H2; g.a_cs = 1;
    g.B = m->W16(2); g.b_rs = 1; g.b_cs = m->H2;
    g.bias = m->Bv(2);
    g.M = B; g.N = m->C; g.K = m->H2;
    uint32_t *lo = lists + size_t(r0) * P, *co = count + r0;
    const dim3 grid(cdiv(B, kToppRows)), blk(256);
    const int KS = (int)m->H2 / 32;
#define RBC_TOPP(ks, pp) hipLaunchKernelGGL((gemm_f16_topp<ks, pp>), grid, blk, 0, st, g, (int)P, tau, lo, co)
#define RBC_TOPP_P(ks) \
  if (P <= 4) RBC_TOPP(ks, 4); else if (P <= 8) RBC_TOPP(ks, 8); else RBC_TOPP(ks, 16)
    switch (KS) {
      case 1: RBC_TOPP_P(1); break;
      case 2: RBC_TOPP_P(2); break;
      case 3: RBC_TOPP_P(3); break;
      case 4: RBC_TOPP_P(4); break;
      case 5: RBC_TOPP_P(5); break;
      case 6: RBC_TOPP_P(6); break;
      case 7: RBC_TOPP_P(7); break;
      default: RBC_TOPP_P(8); break;
    }
#undef RBC_TOPP_P
#undef RBC_TOPP
    RBC_HIP_TRY(hipGetLastError());
  }
  return RBC_OK;
}

int mlp_dims(const rbc_mlp *m, uint32_t *d_in, uint32_t *n_out) {
  *d_in = m->D;
  *n_out = m->C;
  return RBC_OK;
}

}  // namespace rbc

using namespace rbc;

extern "C" {

int rbc_mlp_create(uint32_t d_in, uint32_t h1, uint32_t h2, uint32_t n_out, uint64_t seed, rbc_mlp **out) {
  RBC_ARG_CHECK(out, "NULL output");
  RBC_ARG_CHECK(d_in && h1 && h2 && n_out, "zero layer width");
  auto m = new rbc_mlp();
  m->D = d_in;
  m->H1 = h1;
  m->H2 = h2;
  m->C = n_out;
  size_t sz[6] = {size_t(h1) * d_in, h1, size_t(h2) * h1, h2, size_t(n_out) * h2, n_out};
  size_t o = 0;
  for (int i = 0; i < 6; ++i) {
    m->off[i] = o;
    o += sz[i];
  }
  m->off[6] = o;
  m->np = o;
  std::vector<float> hw(o, 0.f);
  std::mt19937_64 rng(seed);
  const uint32_t fan_in[3] = {d_in, h1, h2};
  for (int layer = 0; layer < 3; ++layer) {
    float a = 1.0f / std::sqrt((float)fan_in[layer]);
    std::uniform_real_distribution<float> U(-a, a);
    for (size_t i = m->off[2 * layer]; i < m->off[2 * layer + 1]; ++i) hw[i] = U(rng);
  }
  int st;
  if ((st = m->w.alloc(o)) || (st = m->g.alloc(o)) || (st = m->m1.alloc(o)) || (st = m->m2.alloc(o)) ||
      (st = m->w16.alloc(o))) {
    delete m;
    return st;
  }
  if (hipMemcpy(m->w.get(), hw.data(), o * 4, hipMemcpyHostToDevice) != hipSuccess) {
    delete m;
    set_error("parameter upload failed");
    return RBC_ERR_HIP;
  }
  (void)fill_f32(m->m1.get(), 0.f, o, nullptr);
  (void)fill_f32(m->m2.get(), 0.f, o, nullptr);
  hipLaunchKernelGGL(to_half, dim3((uint32_t)cdivz(o, 256)), dim3(256), 0, nullptr, m->w.get(), (int64_t)o,
                     m->w16.get());
  if (hipDeviceSynchronize() != hipSuccess) {
    delete m;
    set_error("MLP init failed");
    return RBC_ERR_HIP;
  }
  *out = m;
  return RBC_OK;
}

int rbc_mlp_destroy(rbc_mlp *m) {
  delete m;
  return RBC_OK;
}

int rbc_mlp_num_params(const rbc_mlp *m, size_t *count) {
  RBC_ARG_CHECK(m && count, "NULL argument");
  *count = m->np;
  return RBC_OK;
}

int rbc_mlp_get_params(const rbc_mlp *m, float *params) {
  RBC_ARG_CHECK(m && params, "NULL argument");
  RBC_HIP_TRY(hipMemcpy(params, m->w.get(), m->np * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

// File layout: the 8-byte magic "RBCMLP01", uint32 {d_in, h1, h2, n_out}, uint64 Adam step,
// then the FP32 parameters and the two Adam moment vectors (num_params each).
int rbc_mlp_save(const rbc_mlp *m, const char *path) {
  RBC_ARG_CHECK(m && path, "NULL argument");
  FILE *f = fopen(path, "wb");
  RBC_ARG_CHECK(f, "cannot open %s for writing", path);
  static const char magic[8] = {'R', 'B', 'C', 'M', 'L', 'P', '0', '1'};
  const uint32_t dims[4] = {m->D, m->H1, m->H2, m->C};
  const uint64_t step = m->step;
  std::vector<float> hv(m->np);
  bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(dims, 4, 4, f) == 4 && fwrite(&step, 8, 1, f) == 1;
  const float *src[3] = {m->w.get(), m->m1.get(), m->m2.get()};
  for (int i = 0; ok && i < 3; ++i) {
    if (hipMemcpy(hv.data(), src[i], m->np * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      fclose(f);
      set_error("parameter download failed");
      return RBC_ERR_HIP;
    }
    ok = fwrite(hv.data(), 4, m->np, f) == m->np;
  }
  ok = (fclose(f) == 0) && ok;
  RBC_ARG_CHECK(ok, "write to %s failed", path);
  return RBC_OK;
}

int rbc_mlp_load(const char *path, rbc_mlp **out) {
  RBC_ARG_CHECK(path && out, "NULL argument");
  *out = nullptr;
  FILE *f = fopen(path, "rb");
  RBC_ARG_CHECK(f, "cannot open %s", path);
  char magic[8];
  uint32_t dims[4];
  uint64_t step = 0;
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "RBCMLP01", 8) || fread(dims, 4, 4, f) != 4 ||
      fread(&step, 8, 1, f) != 1) {
    fclose(f);
    set_error("%s is not a proposal-network file", path);
    return RBC_ERR_ARG;
  }
  rbc_mlp *m = nullptr;
  int st = rbc_mlp_create(dims[0], dims[1], dims[2], dims[3], 0, &m);
  if (st != RBC_OK) {
    fclose(f);
    return st;
  }
  std::vector<float> hv(m->np);
  float *dst[3] = {m->w.get(), m->m1.get(), m->m2.get()};
  for (int i = 0; i < 3; ++i) {
    if (fread(hv.data(), 4, m->np, f) != m->np) {
      fclose(f);
      delete m;
      set_error("%s is truncated", path);
      return RBC_ERR_ARG;
    }
    if (hipMemcpy(dst[i], hv.data(), m->np * 4, hipMemcpyHostToDevice) != hipSuccess) {
      fclose(f);
      delete m;
      set_error("parameter upload failed");
      return RBC_ERR_HIP;
    }
  }
  fclose(f);
  m->step = step;
  hipLaunchKernelGGL(to_half, dim3((uint32_t)cdivz(m->np, 256)), dim3(256), 0, nullptr, m->w.get(),
                     (int64_t)m->np, m->w16.get());
  if (hipDeviceSynchronize() != hipSuccess) {
    delete m;
    set_error("proposal network load failed");
    return RBC_ERR_HIP;
  }
  *out = m;
  return RBC_OK;
}

int rbc_mlp_set_params(rbc_mlp *m, const float *params) {
  RBC_ARG_CHECK(m && params, "NULL argument");
  RBC_HIP_TRY(hipMemcpy(m->w.get(), params, m->np * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(to_half, dim3((uint32_t)cdivz(m->np, 256)), dim3(256), 0, nullptr, m->w.get(),
                     (int64_t)m->np, m->w16.get());
  RBC_HIP_TRY(hipDeviceSynchronize());
  return RBC_OK;
}

int rbc_mlp_loss_grad_multi(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, const uint32_t *labels,
                            uint32_t labels_per_row, float *loss, float *grads) {
  RBC_ARG_CHECK(m && x && labels && rows, "bad argument");
  RBC_ARG_CHECK(ldx >= m->D, "ldx < input width");
  RBC_ARG_CHECK(labels_per_row >= 1 && labels_per_row <= 64, "labels per row %u outside [1, 64]",
                labels_per_row);
  hipStream_t st = nullptr;
  RBC_TRY(m->ensure(rows));
  RBC_TRY(m->ensure_labels(size_t(rows) * labels_per_row));
  DevBuf<float> xd;
  RBC_TRY(xd.alloc(size_t(rows) * ldx));
  RBC_HIP_TRY(hipMemcpy(xd.get(), x, size_t(rows) * ldx * 4, hipMemcpyHostToDevice));
  RBC_HIP_TRY(hipMemcpy(m->lab.get(), labels, size_t(rows) * labels_per_row * 4, hipMemcpyHostToDevice));
  std::vector<uint32_t> hi(rows);
  for (uint32_t i = 0; i < rows; ++i) hi[i] = i;
  RBC_HIP_TRY(hipMemcpy(m->idx.get(), hi.data(), size_t(rows) * 4, hipMemcpyHostToDevice));
  int64_t ne = int64_t(rows) * m->D;
  hipLaunchKernelGGL(gather_half, dim3((uint32_t)cdivz(ne, 256)), dim3(256), 0, st, xd.get(), (int64_t)ldx,
                     m->idx.get(), (int)rows, (int)m->D, m->x16.get());
  RBC_TRY(mlp_forward(m, rows, nullptr, st));
  RBC_TRY(mlp_backward(m, rows, labels_per_row, st));
  RBC_HIP_TRY(hipDeviceSynchronize());
  if (loss) RBC_HIP_TRY(hipMemcpy(loss, m->loss.get(), 4, hipMemcpyDeviceToHost));
  if (grads) RBC_HIP_TRY(hipMemcpy(grads, m->g.get(), m->np * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

int rbc_mlp_loss_grad(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, const uint32_t *labels,
                      float *loss, float *grads) {
  return rbc_mlp_loss_grad_multi(m, x, rows, ldx, labels, 1, loss, grads);
}

int rbc_mlp_logits(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, float *logits) {
  RBC_ARG_CHECK(m && x && logits && rows, "bad argument");
  RBC_TRY(m->ensure(rows));
  DevBuf<float> xd, zd;
  RBC_TRY(xd.alloc(size_t(rows) * ldx));
  RBC_TRY(zd.alloc(size_t(rows) * m->C));
  RBC_HIP_TRY(hipMemcpy(xd.get(), x, size_t(rows) * ldx * 4, hipMemcpyHostToDevice));
  std::vector<uint32_t> hi(rows);
  for (uint32_t i = 0; i < rows; ++i) hi[i] = i;
  RBC_HIP_TRY(hipMemcpy(m->idx.get(), hi.data(), size_t(rows) * 4, hipMemcpyHostToDevice));
  int64_t ne = int64_t(rows) * m->D;
  hipLaunchKernelGGL(gather_half, dim3((uint32_t)cdivz(ne, 256)), dim3(256), 0, nullptr, xd.get(), (int64_t)ldx,
                     m->idx.get(), (int)rows, (int)m->D, m->x16.get());
  RBC_TRY(mlp_forward(m, rows, zd.get(), nullptr));
  RBC_HIP_TRY(hipDeviceSynchronize());
  RBC_HIP_TRY(hipMemcpy(logits, zd.get(), size_t(rows) * m->C * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

int rbc_mlp_propose(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, uint32_t *out) {
  RBC_ARG_CHECK(m && x && out && rows, "bad argument");
  RBC_ARG_CHECK(ldx >= m->D, "ldx < input width");
  DevBuf<float> xd;
  DevBuf<uint32_t> od;
  RBC_TRY(xd.alloc(size_t(rows) * ldx));
  RBC_TRY(od.alloc(rows));
  RBC_HIP_TRY(hipMemcpy(xd.get(), x, size_t(rows) * ldx * 4, hipMemcpyHostToDevice));
  RBC_TRY(mlp_propose(m, xd.get(), ldx, rows, od.get(), nullptr));
  RBC_HIP_TRY(hipMemcpy(out, od.get(), size_t(rows) * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

int rbc_mlp_propose_lists(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, uint32_t P, float mass,
                          uint32_t *lists, uint32_t *counts) {
  RBC_ARG_CHECK(m && x && lists && counts && rows, "bad argument");
  RBC_ARG_CHECK(ldx >= m->D, "ldx < input width");
  DevBuf<float> xd;
  DevBuf<uint32_t> ld, cd;
  RBC_TRY(xd.alloc(size_t(rows) * ldx));
  RBC_TRY(ld.alloc(size_t(rows) * P));
  RBC_TRY(cd.alloc(rows));
  RBC_HIP_TRY(hipMemcpy(xd.get(), x, size_t(rows) * ldx * 4, hipMemcpyHostToDevice));
  RBC_TRY(mlp_propose_lists(m, xd.get(), ldx, rows, P, mass, ld.get(), cd.get(), nullptr));
  RBC_HIP_TRY(hipMemcpy(lists, ld.get(), size_t(rows) * P * 4, hipMemcpyDeviceToHost));
  RBC_HIP_TRY(hipMemcpy(counts, cd.get(), size_t(rows) * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

int rbc_mlp_train_multi(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, const uint32_t *labels,
                        uint32_t labels_per_row, int mem, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                        float *epoch_loss) {
  RBC_ARG_CHECK(m && x && labels && rows && batch, "bad argument");
  RBC_ARG_CHECK(ldx >= m->D, "ldx < input width");
  const uint32_t L = labels_per_row;
  RBC_ARG_CHECK(L >= 1 && L <= 64, "labels per row %u outside [1, 64]", L);
  hipStream_t st = nullptr;
  RBC_TRY(m->ensure(batch));
  RBC_TRY(m->ensure_labels(size_t(batch) * L));
  DevMat X;
  RBC_TRY(X.stage(x, rows, m->D, ldx, mem, st));
  DevBuf<uint32_t> labd, perm;
  RBC_TRY(labd.alloc(size_t(rows) * L));
  RBC_TRY(perm.alloc(rows));
  RBC_TRY(copy_in(labd.get(), labels, size_t(rows) * L * 4, mem, st));
  std::vector<uint32_t> hp(rows);
  for (uint32_t i = 0; i < rows; ++i) hp[i] = i;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    std::shuffle(hp.begin(), hp.end(), rng);
    // Synchronous upload: the next epoch's shuffle rewrites hp, so the copy must be done
    // before this loop moves on (the previous epoch's kernels that read perm are ordered
    // before it on the stream).
    RBC_TRY(upload_host(perm.get(), hp.data(), size_t(rows) * 4, st));
    uint32_t nb = 0;
    for (uint32_t r0 = 0; r0 + batch <= rows; r0 += batch, ++nb) {
      const uint32_t B = batch;
      int64_t ne = int64_t(B) * m->D;
      hipLaunchKernelGGL(gather_half, dim3((uint32_t)cdivz(ne, 256)), dim3(256), 0, st, X.p, (int64_t)X.ld,
                         perm.get() + r0, (int)B, (int)m->D, m->x16.get());
      hipLaunchKernelGGL(gather_u32, dim3(cdiv(B * L, 256)), dim3(256), 0, st, labd.get(), perm.get() + r0,
                         (int)B, (int)L, m->lab.get());
      RBC_HIP_TRY(hipGetLastError());
      RBC_TRY(mlp_forward(m, B, nullptr, st));
      RBC_TRY(mlp_backward(m, B, L, st));
      ++m->step;
      float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
      hipLaunchKernelGGL(adam_step, dim3((uint32_t)cdivz(m->np, 256)), dim3(256), 0, st, m->w.get(), m->w16.get(),
                         m->g.get(), m->m1.get(), m->m2.get(), (int64_t)m->np, lr, b1, b2, eps, c1, c2);
      RBC_HIP_TRY(hipGetLastError());
    }
    if (epoch_loss) {
      RBC_HIP_TRY(hipStreamSynchronize(st));
      float l = 0.f;
      RBC_HIP_TRY(hipMemcpy(&l, m->loss.get(), 4, hipMemcpyDeviceToHost));
      epoch_loss[ep] = nb ? l : 0.f;  // loss of the epoch's last batch
    }
  }
  RBC_HIP_TRY(hipDeviceSynchronize());
  return RBC_OK;
}

int rbc_mlp_train(rbc_mlp *m, const float *x, uint32_t rows, uint32_t ldx, const uint32_t *labels, int mem,
                  uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *epoch_loss) {
  return rbc_mlp_train_multi(m, x, rows, ldx, labels, 1, mem, epochs, batch, lr, seed, epoch_loss);
}

}  // extern "C"
