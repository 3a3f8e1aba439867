// Batched exact k-NN: many independent exact-RBC indexes, each with its own query batch,
// answered by one persistent kernel launch (captured in a hipGraph and replayed).
//
// Work unit = a tile of up to QT queries of one index.  Persistent workgroups pull tiles
// from a global counter across all indexes.  For its tile a workgroup
//   1. computes every query's FP32 distance to every representative of the index (kept
//      in a per-workgroup scratch row block),
//   2. bounds each list from below for the whole tile (the minimum over the tile's
//      queries of the triangle/Voronoi lower bound) and sorts the lists by it in LDS,
//   3. scans lists in that order with the FP32 tile engine, keeping every query's best K'
//      in LDS, skipping a list when no query of the tile can still find a point in it
//      and stopping when the next list's bound exceeds every query's current k-th
//      distance (the bound tightens as lists are scanned),
//   4. writes each query's K' candidates and whether they are provably complete.
// A second kernel re-ranks the candidates exactly (tree-ordered FP32, FP64 near-tie
// guard) and writes each index's results.  Queries whose K' could not be proven
// complete are re-run through the single-index pipeline.
//
// The tile size QT is picked by a small autotuner: the first query call on a batch shape
// times QT = 16, 32 and 64 and keeps the fastest.
#include <algorithm>
#include <map>
#include <vector>

#include "exact_device.hpp"
#include "exact_query.hpp"

struct rbc_batched;

namespace rbc {

namespace {

struct BInst {
  const float *xp;
  const float *reps;
  const uint32_t *list_off;
  const uint32_t *perm;
  const uint32_t *inv;
  const float *radius;
  const float *rr;
  const float *xn2;  // ||x||^2 per permuted row (FP64 sum rounded to nearest)
  const uint16_t *rhi, *rlo;  // the reps split into BF16 hi/lo rows (ldb)
  const float *rn2;           // ||r||^2
  uint32_t ldb;
  float xmax;        // max ||x|| over the index (rounded up)
  double csplit;     // split-BF16 dot-product constant for this index's width
  uint32_t n, d, dpad, ld, R, P;  // P: tree width (64, 128)
  const float *q;                 // this index's queries (ldq), m of them
  uint32_t ldq, m, qbase;         // qbase: first global query number
  float *out_d;
  uint32_t *out_i;
  uint32_t ldo;
};

struct BTile {
  uint32_t inst, q0, qn;
};

struct BatchArgs {
  const BInst *inst;
  const BTile *tiles;
  uint32_t ntiles;
  uint32_t *next;     // work counter
  float *scratch;     // per workgroup: kTQ x Rmax
  uint32_t Rmax, Rpad;
  uint32_t k, Kp;
  double e;           // FP32 FMA-order relative error
  int metric;
  uint32_t *cand_i;   // [mtot][Kp]
  float *cand_d;      // [mtot][Kp]
  uint32_t *cert;     // [mtot]
  uint32_t *qinst;    // [mtot] index of each global query
  unsigned long long *stats;  // [4] (query, point) pairs scanned, [5] their FLOPs + the q->rep FLOPs,
                              // [6] 64-point list steps, [7] longest tile (wall-clock ticks)
};

// Dense distances of the tile's queries to a block of reps into the scratch rows.
struct EpiScratch {
  float *out;
  uint32_t ld;
  uint32_t col0;
  static __host__ __device__ size_t lds_bytes() { return 0; }
  __device__ void begin(const TileCtx &) {}
  __device__ void step(const TileCtx &c, float (&acc)[4][4], uint32_t p0) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        uint32_t p = p0 + 4 * c.pg + b;
        if (c.pid_s[4 * c.pg + b] != kDummy) out[(4 * c.qg + a) * ld + col0 + p] = acc[a][b];
      }
  }
  __device__ void end(const TileCtx &) {}
};

__host__ __device__ inline size_t batch_lds_bytes(uint32_t dpad, uint32_t Rpad) {
  return tile_lds_base_bytes(dpad) + EpiTopK::lds_bytes() + size_t(Rpad) * 8 + kTQ * 16 + 64 + kTQ * 8;
}

// ---- split-BF16 MFMA list scan (L2, d <= 128) ----------------------------------------
// The tile's 64 queries are MFMA A operands (wave w: queries 16w.., 16 rows), kept in
// registers for the whole tile as hi/lo BF16 pairs; a list is scanned 64 points per step,
// four 16-point B tiles per wave, each split into hi/lo from its FP32 row on the fly.
// q.x ~ qh.xh + qh.xl + ql.xh (three v_mfma_f32_16x16x32_bf16 per k-step) and
// D~ = ||q||^2 + ||x||^2 - 2 q.x, within E_q of the true value for every point of the
// index (the bound of the query-rep kernel with the index's largest point norm).  Lane l
// holds queries 16w + 4(l/16) + i of point 16t + l%16; candidates below a query's running
// K'-th value are staged and folded exactly as EpiTopK does (on D~ instead of FP32 values).
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(8 * sizeof(uint16_t)))) uint16_t u16x8;

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// hi = bf16(v), lo = bf16(v - hi) of 8 coordinates.
__device__ __forceinline__ void split8(const float (&v)[8], bf16x8 &hi, bf16x8 &lo) {
  u16x8 h, o;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const uint16_t b = bf16_rne(v[e]);
    h[e] = b;
    o[e] = bf16_rne(v[e] - __uint_as_float(uint32_t(b) << 16));
  }
  hi = __builtin_bit_cast(bf16x8, h);
  lo = __builtin_bit_cast(bf16x8, o);
}

// Coordinates c0..c0+7 of a row of d floats (zero from d on).
__device__ __forceinline__ void load8(const float *row, uint32_t c0, uint32_t d, bool live, float (&v)[8]) {
  if (live && c0 + 8 <= d && (reinterpret_cast<uintptr_t>(row + c0) & 15u) == 0) {
    const float4 a = *reinterpret_cast<const float4 *>(row + c0), b = *reinterpret_cast<const float4 *>(row + c0 + 4);
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = live && c0 + e < d ? row[c0 + e] : 0.f;
  }
}

// Stages (v, id) for query qi when it beats the query's running K'-th pair (thd, thi: the
// pair as of the last fold; EpiTopK::step's first half).
__device__ __forceinline__ void topk_stage(const EpiTopK::L &L, uint32_t qi, float v, uint32_t id, float thd,
                                           uint32_t thi) {
  if (id != kDummy && pair_less(v, id, thd, thi)) {
    const uint32_t s = atomicAdd(&L.cnt[qi], 1u);
    L.cd[qi * 64 + s] = v;
    L.ci[qi * 64 + s] = id;
  }
}

// One thread per query folds its staged pairs into its sorted K best (EpiTopK::step's
// second half); brackets itself with barriers.
__device__ void topk_fold(const EpiTopK::L &L, uint32_t K) {
  __syncthreads();
  if (threadIdx.x < kTQ) {
    const int qi = threadIdx.x;
    const uint32_t n = L.cnt[qi];
    if (n) {
      float kd = L.thd[qi];
      uint32_t ki = L.thi[qi];
      float *td = L.td + qi * 64;
      uint32_t *ti = L.ti + qi * 64;
      for (uint32_t c = 0; c < n; ++c) {
        const float cv = L.cd[qi * 64 + c];
        const uint32_t cid = L.ci[qi * 64 + c];
        if (!pair_less(cv, cid, kd, ki)) continue;
        int pos = (int)K - 1;
        while (pos > 0 && pair_less(cv, cid, td[pos - 1], ti[pos - 1])) {
          td[pos] = td[pos - 1];
          ti[pos] = ti[pos - 1];
          --pos;
        }
        td[pos] = cv;
        ti[pos] = cid;
        kd = td[K - 1];
        ki = ti[K - 1];
      }
      L.thd[qi] = kd;
      L.thi[qi] = ki;
      L.cnt[qi] = 0;
    }
  }
  __syncthreads();
}

// Scans permuted rows [off, off + len) of index I for the tile's queries (Qh/Ql: this
// wave's A operands, qn2s: their ||q||^2, nq: live queries).  The kernel runs one
// workgroup per CU (its LDS), so each wave has the registers to hold the next step's point
// rows: they are loaded while the current step's MFMAs, staging and fold run.
__device__ void mfma_scan_list(const BInst &I, uint32_t off, uint32_t len, const bf16x8 (&Qh)[4],
                               const bf16x8 (&Ql)[4], const float *qn2s, uint32_t nq, const EpiTopK::L &L,
                               uint32_t K) {
  const int w = threadIdx.x / kWave, l = lane_id();
  const uint32_t koff = (l / 16) * 8;
  const int KS = (int)((I.d + 31u) / 32u);
  float nv[4][4][8];  // next step: [tile][k-step][8 coordinates]
  uint32_t nid[4];
  float nxn[4];
  auto fetch = [&](uint32_t p0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t p = p0 + 16 * t + (l % 16);
      const bool live = p < len;
      const uint32_t row = off + (live ? p : 0);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) load8(I.xp + size_t(row) * I.ld, ks * 32 + koff, I.d, live && ks < KS, nv[t][ks]);
      nid[t] = live ? I.perm[row] : kDummy;
      nxn[t] = live ? I.xn2[row] : 0.f;
    }
  };
  fetch(0);
  for (uint32_t p0 = 0; p0 < len; p0 += kTP) {
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      if (ks >= KS) break;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        bf16x8 bh, bl;
        split8(nv[t][ks], bh, bl);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Qh[ks], bh, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Qh[ks], bl, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ql[ks], bh, acc[t], 0, 0, 0);
      }
    }
    uint32_t id[4];
    float xn[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      id[t] = nid[t];
      xn[t] = nxn[t];
    }
    // The next step's rows load during this step's staging and fold.
    if (p0 + kTP < len) fetch(p0 + kTP);
    // This lane's four queries: their norms and running K'-th pairs (changed only by folds).
    float qn[4], thd[4];
    uint32_t thi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t qi = 16 * w + 4 * (l / 16) + i;
      qn[i] = qn2s[qi];
      thd[i] = L.thd[qi];
      thi[i] = L.thi[qi];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t qi = 16 * w + 4 * (l / 16) + i;
        if (qi >= nq) continue;
        const float v = fmaxf(__builtin_fmaf(-2.f, acc[t][i], qn[i] + xn[t]), 0.f);
        topk_stage(L, qi, v, id[t], thd[i], thi[i]);
      }
    }
    topk_fold(L, K);
  }
}

template <int METRIC>
__global__ __launch_bounds__(kNT) void batched_exact_kernel(BatchArgs a, uint32_t dpad_max) {
  extern __shared__ float smem[];
  const int t = threadIdx.x;
  // LDS: [tile engine + TopK state][tl: Rpad floats][ord: Rpad u32][j0, d0, gam per query][misc]
  unsigned char *base = reinterpret_cast<unsigned char *>(smem);
  float *tl = reinterpret_cast<float *>(base + tile_lds_base_bytes(dpad_max) + EpiTopK::lds_bytes());
  uint32_t *ord = reinterpret_cast<uint32_t *>(tl + a.Rpad);
  uint32_t *j0s = ord + a.Rpad;
  float *d0s = reinterpret_cast<float *>(j0s + kTQ);
  float *gam = d0s + kTQ;
  float *gtmp = gam + kTQ;  // [kTQ] scratch
  uint32_t *misc = reinterpret_cast<uint32_t *>(gtmp + kTQ);  // [0] tile, [1] need flag, [2] gmax bits
  float *qn2s = reinterpret_cast<float *>(misc + 16);          // [kTQ] ||q||^2 (MFMA scan)
  float *eqs = qn2s + kTQ;                                     // [kTQ] absolute error of D~
  float *scr = a.scratch + size_t(blockIdx.x) * kTQ * a.Rmax;

  for (;;) {
    __syncthreads();
    if (t == 0) misc[0] = atomicAdd(a.next, 1u);
    __syncthreads();
    const uint32_t tile = misc[0];
    if (tile >= a.ntiles) break;
    const BTile T = a.tiles[tile];
    const unsigned long long t_begin = wall_clock64();
    unsigned long long tile_pairs = 0, tile_steps_n = 0;  // thread 0's per-tile counts
    const BInst I = a.inst[T.inst];
    ScanItem it{0, 0, T.q0, T.qn, T.inst};
    // 1. queries -> LDS; distances to all reps -> scratch.
    TileSrc qsrc{I.xp, I.ld, nullptr, I.perm, I.q, I.ldq, nullptr, I.d, I.dpad};
    TileCtx ctx = tile_setup(qsrc, it, smem);
    {
      TileSrc rsrc{I.reps, I.ld, nullptr, nullptr, I.q, I.ldq, nullptr, I.d, I.dpad};
      EpiScratch es{scr, a.Rmax, 0};
      tile_steps<METRIC>(rsrc, 0, I.R, es, ctx, smem);
      if (t == 0) atomicAdd(&a.stats[5], 2ull * I.d * T.qn * I.R);
    }
    __syncthreads();
    // Nearest rep per query (lowest index on ties).
    if (t < kTQ) {
      float bv = FLT_MAX;
      uint32_t bj = 0;
      if ((uint32_t)t < T.qn)
        for (uint32_t j = 0; j < I.R; ++j) {
          float v = scr[t * a.Rmax + j];
          if (v < bv) {
            bv = v;
            bj = j;
          }
        }
      j0s[t] = bj;
      d0s[t] = bv;
      gam[t] = INFINITY;
    }
    __syncthreads();
    // 2. Tile lower bound of every list, then sort ascending.
    for (uint32_t j = t; j < a.Rpad; j += kNT) {
      float b = INFINITY;
      if (j < I.R) {
        double psi = (double)I.radius[j];
        double mn = INFINITY;
        for (uint32_t qi = 0; qi < T.qn; ++qi) {
          double lb = list_lower_bound((double)scr[qi * a.Rmax + j], (double)d0s[qi],
                                       (double)I.rr[size_t(j0s[qi]) * I.R + j], psi, a.e, j == j0s[qi], METRIC);
          mn = fmin(mn, lb);
        }
        b = (float)mn;
        if ((double)b > mn) b = nextafterf(b, -INFINITY);
      }
      tl[j] = b;
      ord[j] = j;
    }
    for (uint32_t kk = 2; kk <= a.Rpad; kk <<= 1)
      for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
        __syncthreads();
        for (uint32_t i = t; i < a.Rpad; i += kNT) {
          uint32_t p = i ^ jj;
          if (p > i) {
            bool up = (i & kk) == 0;
            float vi = tl[i], vp = tl[p];
            bool sw = up ? (vp < vi || (vp == vi && ord[p] < ord[i])) : (vi < vp || (vi == vp && ord[i] < ord[p]));
            if (sw) {
              tl[i] = vp;
              tl[p] = vi;
              uint32_t o = ord[i];
              ord[i] = ord[p];
              ord[p] = o;
            }
          }
        }
      }
    __syncthreads();
    // 3. Scan lists in bound order with per-query best K' (FP32 tiles, or split-BF16 MFMA
    // tiles for L2 with values D~ within eqs[q] of the true distance).
    const bool mf = METRIC == 0 && I.d <= 128;
    bf16x8 Qh[4], Ql[4];
    if (mf) {
      const int w = t / kWave, l = lane_id();
      const uint32_t qi = 16 * w + (l % 16);
      const bool ok = qi < T.qn;
      const float *qr = I.q + size_t(T.q0 + (ok ? qi : 0)) * I.ldq;
      double s2 = 0.0;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        float v[8];
        load8(qr, ks * 32 + (l / 16) * 8, I.d, ok && (uint32_t)ks * 32 < I.d, v);
#pragma unroll
        for (int e = 0; e < 8; ++e) s2 += (double)v[e] * (double)v[e];
        split8(v, Qh[ks], Ql[ks]);
      }
      s2 += xor_lane(s2, 16);
      s2 += xor_lane(s2, 32);
      if (l < 16 && ok) {
        qn2s[qi] = (float)s2;
        const double u = 0x1p-24, nq = sqrt(s2), nx = (double)I.xmax, c = I.csplit;
        const double E = (2.0 * c * nq * nx + 4.0 * u * (s2 + nx * nx) + 2.0 * u * (1.0 + c) * nq * nx) * 1.01 + 1e-30;
        float Ef = (float)E;
        if ((double)Ef < E) Ef = nextafterf(Ef, FLT_MAX);
        eqs[qi] = Ef;
      }
    }
    EpiTopK epi{a.Kp, nullptr, nullptr, 0, 0};
    epi.begin(ctx);
    if (t == 0) misc[2] = __float_as_uint(INFINITY);
    __syncthreads();
    for (uint32_t idx = 0; idx < I.R; ++idx) {
      __syncthreads();  // everyone has read the previous iteration's flag
      const float gmax = __uint_as_float(misc[2]);
      if (tl[idx] > gmax) break;  // every later list is out of every query's reach
      const uint32_t j = ord[idx];
      if (t == 0) misc[1] = 0;
      __syncthreads();
      if ((uint32_t)t < T.qn) {
        double lb = list_lower_bound((double)scr[t * a.Rmax + j], (double)d0s[t],
                                     (double)I.rr[size_t(j0s[t]) * I.R + j], (double)I.radius[j], a.e,
                                     j == j0s[t], METRIC);
        if (lb <= (double)gam[t]) misc[1] = 1;
      }
      __syncthreads();
      if (!misc[1]) continue;
      const uint32_t len = I.list_off[j + 1] - I.list_off[j];
      if (mf)
        mfma_scan_list(I, I.list_off[j], len, Qh, Ql, qn2s, T.qn, epi.lay(ctx), a.Kp);
      else
        tile_steps<METRIC>(qsrc, I.list_off[j], len, epi, ctx, smem);
      tile_pairs += (unsigned long long)T.qn * len;
      tile_steps_n += (len + kTP - 1) / kTP;
      // Refresh each query's bound from its current k-th value.
      EpiTopK::L L = epi.lay(ctx);
      if (t < kTQ) {
        float g = INFINITY;
        if ((uint32_t)t < T.qn) {
          float v = L.td[t * 64 + (a.k - 1)];
          if (v < FLT_MAX) {
            double ghi = mf ? ((double)v + (double)eqs[t]) * (1.0 + 1e-12) : (double)v / (1.0 - a.e) * (1.0 + 1e-12);
            double gd = METRIC == 0 ? sqrt(ghi) * (1.0 + 1e-12) : ghi;
            g = (float)gd;
            if ((double)g < gd) g = nextafterf(g, INFINITY);
          }
          gam[t] = g;
        }
        gtmp[t] = (uint32_t)t < T.qn ? g : 0.f;
      }
      // Wave 0 (the threads t < kTQ that wrote them) reduces the bounds to their maximum.
      if (t < kTQ) {
        float mx = gtmp[t];
#pragma unroll
        for (int s = kWave / 2; s > 0; s >>= 1) mx = fmaxf(mx, xor_lane(mx, s));
        if (t == 0) misc[2] = __float_as_uint(mx);
      }
      __syncthreads();
    }
    if (t == 0) {
      atomicAdd(&a.stats[4], tile_pairs);
      atomicAdd(&a.stats[5], 2ull * I.d * tile_pairs);
      atomicAdd(&a.stats[6], tile_steps_n);
      atomicMax(&a.stats[7], wall_clock64() - t_begin);
    }
    // 4. Candidates and completeness.
    EpiTopK::L L = epi.lay(ctx);
    for (uint32_t i = t; i < kTQ * a.Kp; i += kNT) {
      const uint32_t qi = i / a.Kp, jj = i % a.Kp;
      if (qi >= T.qn) continue;
      const size_t g = size_t(I.qbase + T.q0 + qi);
      a.cand_i[g * a.Kp + jj] = L.ti[qi * 64 + jj];
      a.cand_d[g * a.Kp + jj] = L.td[qi * 64 + jj];
      if (jj == 0) {
        const float vk = L.td[qi * 64 + a.k - 1], vK = L.td[qi * 64 + a.Kp - 1];
        // Points left out have a value >= vK; they cannot beat the k-th kept when
        // vK / (1 + e) > vk / (1 - e) (FP32 values), or vK - E > vk + E (D~).
        bool ok = vK >= FLT_MAX ||
                  (mf ? (double)vK - (double)eqs[qi] > ((double)vk + (double)eqs[qi]) * (1.0 + 1e-12)
                      : (double)vK / (1.0 + a.e) > (double)vk / (1.0 - a.e) * (1.0 + 1e-12));
        a.cert[g] = ok ? 1u : 0u;
      }
    }
  }
}

struct BSelArgs {
  const BInst *inst;
  const uint32_t *qinst;
  uint32_t mtot, k, Kp, P;
  const uint32_t *cand_i;
  const uint32_t *cert;
  unsigned long long *stats;
};

template <int METRIC, int NPL>
__global__ __launch_bounds__(256) void batched_select(BSelArgs a) {
  const uint32_t g = blockIdx.x * 4 + threadIdx.x / kWave;
  if (g >= a.mtot) return;
  const BInst &I = a.inst[a.qinst[g]];
  if (I.P != a.P || !a.cert[g]) return;
  const int l = lane_id();
  const uint32_t q = g - I.qbase;
  const float *qrow = I.q + size_t(q) * I.ldq;
  float qv[NPL];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    uint32_t c = l + kWave * j;
    qv[j] = c < I.d ? qrow[c] : 0.f;
  }
  float tv = FLT_MAX;
  uint32_t ti = kDummy;
  auto cand = [&](uint32_t e) -> uint32_t {
    const uint32_t id = a.cand_i[size_t(g) * a.Kp + e];
    return id == kDummy ? id : RBC_CHECKED(id, I.n);
  };
  rerank_candidates<METRIC, NPL>(tv, ti, a.Kp, cand, I.xp, I.ld, I.inv, qv, I.d);
  const double P = (double)a.P;
  const double e_tree = ((P <= 64 ? 6.0 : 7.0) + 4.0) * 0x1p-24 * 1.01;
  near_tie_guard<METRIC>(tv, ti, a.Kp, I.xp, I.ld, I.inv, qrow, I.d, a.k, e_tree, a.stats);
  if (l < (int)a.k) {
    I.out_d[size_t(q) * I.ldo + l] = tv;
    I.out_i[size_t(q) * I.ldo + l] = ti;
  }
}

constexpr int kBatchStats = 8;

// ||x||^2 of every row (FP64 sum rounded to nearest).
__global__ void row_norms2(const float *xp, uint32_t n, uint32_t d, uint32_t ld, float *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (uint32_t c = 0; c < d; ++c) s += (double)xp[size_t(i) * ld + c] * (double)xp[size_t(i) * ld + c];
  out[i] = (float)s;
}

// ---- wave-autonomous batched search (L2, d <= 128, K' <= 32) --------------------------
// Each wave is an independent searcher: it takes a tile of up to 16 queries of one index
// from a global counter and runs the whole exact search for them with no workgroup
// barrier, so a CU holds as many searching waves as its LDS and registers allow.
//   1. The 16 queries become MFMA A operands (hi/lo BF16 in registers), with ||q||^2 and
//      the bound E_q of the split-BF16 distances (largest point norm of the index).
//   2. D~(q, r) for every rep on split-BF16 MFMA tiles (the reps' hi/lo rows), kept in a
//      per-wave global scratch block (16 x Rmax, L2 resident).
//   3. The nearest rep of each query; the lower bound of every list over the wave's
//      queries (triangle with psi and the Voronoi plane against each query's nearest rep,
//      with E_q and the FP32 slack e of rr and of the owner test); the lists sorted by it.
//   4. Lists in that order, skipping a list when no query can still find a point in it and
//      stopping when the next bound exceeds every query's gamma; each list is scanned 64
//      points per step as in mfma_scan_list, candidates under a query's running K'-th pair
//      staged in the wave's LDS and folded by one lane per query.
//   5. The K' candidates and their certificate (D~_K' - E > D~_k + E), as the other kernel.
constexpr uint32_t kWQ = 16;   // queries per wave tile (one MFMA row block)
constexpr uint32_t kWK = 32;   // K' slots per query
constexpr uint32_t kWS = 64;   // staging slots per query (one 64-point step)

struct WaveLds {
  float td[kWQ][kWK];
  uint32_t ti[kWQ][kWK];
  float cd[kWQ][kWS];
  uint32_t ci[kWQ][kWS];
  uint32_t cnt[kWQ];
  float thd[kWQ];
  uint32_t thi[kWQ];
  float gam[kWQ];   // metric-distance bound on each query's k-th NN (INF: none yet)
  float eq[kWQ];    // |D~ - D| bound
  float qn2[kWQ];
  uint32_t j0[kWQ];  // nearest rep
};

__host__ __device__ inline size_t wave_lds_bytes(uint32_t Rpad) {
  return (sizeof(WaveLds) + size_t(Rpad) * 8 + 15) & ~size_t(15);
}

struct WaveArgs {
  const BInst *inst;
  const BTile *tiles;
  uint32_t ntiles;
  uint32_t *next;
  float *scratch;  // per wave: kWQ x Rmax query-rep distances
  uint32_t Rmax, Rpad;
  uint32_t k, Kp;
  double e;        // FP32 relative error of rr and of the owner test
  uint32_t *cand_i;
  float *cand_d;
  uint32_t *cert;
  unsigned long long *stats;
};

// Lower bound (metric distance, L2) on d(q, x) for the points x of list j: the triangle
// with psi_j and the Voronoi plane against the query's nearest rep j0 (D~ within E).
__device__ __forceinline__ double wave_list_bound(double Dqj, double Dq0, double E, double rr_j0, double psi,
                                                  double e, bool same) {
  const double lo = fmax(Dqj - E, 0.0);
  double lb = sqrt(lo) - psi;
  if (!same && rr_j0 > 0.0) {
    const double num = lo - (Dq0 + E) - 2.0 * e * psi * psi;
    const double den = 2.0 * sqrt(rr_j0 * (1.0 + e));
    if (num > 0.0) lb = fmax(lb, num / den);
  }
  return lb * (1.0 - 1e-12);
}

__global__ __launch_bounds__(256, 2) void batched_wave_kernel(WaveArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char wave_smem[];
  const int w = threadIdx.x / kWave, l = lane_id();
  const size_t wl = wave_lds_bytes(a.Rpad);
  WaveLds &S = *reinterpret_cast<WaveLds *>(wave_smem + w * wl);
  float *lbs = reinterpret_cast<float *>(wave_smem + w * wl + sizeof(WaveLds));
  uint32_t *ord = reinterpret_cast<uint32_t *>(lbs + a.Rpad);
  float *dq = a.scratch + size_t(blockIdx.x * (blockDim.x / kWave) + w) * kWQ * a.Rmax;
  const uint32_t koff = (l / 16) * 8;
  for (;;) {
    uint32_t tile = 0;
    if (l == 0) tile = atomicAdd(a.next, 1u);
    tile = __builtin_amdgcn_readfirstlane(tile);
    if (tile >= a.ntiles) break;
    const BTile T = a.tiles[tile];
    const BInst I = a.inst[T.inst];
    const unsigned long long t_begin = wall_clock64();
    const uint32_t nq = T.qn, R = I.R;
    const int KS = (int)((I.d + 31u) / 32u);
    // 1. Queries.
    bf16x8 Qh[4], Ql[4];
    {
      const uint32_t qi = l % 16;
      const bool ok = qi < nq;
      const float *qr = I.q + size_t(T.q0 + (ok ? qi : 0)) * I.ldq;
      double s2 = 0.0;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        float v[8];
        load8(qr, ks * 32 + koff, I.d, ok && ks < KS, v);
#pragma unroll
        for (int e = 0; e < 8; ++e) s2 += (double)v[e] * (double)v[e];
        split8(v, Qh[ks], Ql[ks]);
      }
      s2 += xor_lane(s2, 16);
      s2 += xor_lane(s2, 32);
      if (l < 16) {
        S.qn2[qi] = (float)s2;
        const double u = 0x1p-24, nqn = sqrt(s2), nx = (double)I.xmax, c = I.csplit;
        const double E = (2.0 * c * nqn * nx + 4.0 * u * (s2 + nx * nx) + 2.0 * u * (1.0 + c) * nqn * nx) * 1.01 + 1e-30;
        float Ef = (float)E;
        if ((double)Ef < E) Ef = nextafterf(Ef, FLT_MAX);
        S.eq[qi] = Ef;
        S.gam[qi] = INFINITY;
        S.cnt[qi] = 0;
        S.thd[qi] = FLT_MAX;
        S.thi[qi] = kDummy;
      }
      for (uint32_t i = l; i < kWQ * kWK; i += kWave) {
        (&S.td[0][0])[i] = FLT_MAX;
        (&S.ti[0][0])[i] = kDummy;
      }
    }
    __builtin_amdgcn_wave_barrier();
    float qn4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) qn4[i] = S.qn2[4 * (l / 16) + i];
    // 2. Query-rep distances (reps as 16-column B tiles).
    for (uint32_t r0 = 0; r0 < R; r0 += 64) {
      f32x4 acc[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        if (ks >= KS) break;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const uint32_t r = r0 + 16 * t + (l % 16);
          bf16x8 bh = bf16x8{}, bl = bf16x8{};
          if (r < R) {
            bh = *reinterpret_cast<const bf16x8 *>(I.rhi + size_t(r) * I.ldb + ks * 32 + koff);
            bl = *reinterpret_cast<const bf16x8 *>(I.rlo + size_t(r) * I.ldb + ks * 32 + koff);
          }
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Qh[ks], bh, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Qh[ks], bl, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ql[ks], bh, acc[t], 0, 0, 0);
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t r = r0 + 16 * t + (l % 16);
        if (r >= R) continue;
        const float nr = I.rn2[r];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t q = 4 * (l / 16) + i;
          dq[size_t(q) * a.Rmax + r] = fmaxf(__builtin_fmaf(-2.f, acc[t][i], qn4[i] + nr), 0.f);
        }
      }
    }
    __threadfence_block();  // the distance block is read back by other lanes below
    // 3. Nearest rep per query (four lanes per query, lowest index among equals).
    {
      const uint32_t qi = l % 16;
      float bv = FLT_MAX;
      uint32_t bj = kDummy;
      for (uint32_t j = l / 16; j < R; j += 4) {
        const float v = dq[size_t(qi) * a.Rmax + j];
        if (v < bv) {
          bv = v;
          bj = j;
        }
      }
#pragma unroll
      for (int s = 16; s <= 32; s <<= 1) {
        const float ov = xor_lane(bv, s);
        const uint32_t oj = xor_lane(bj, s);
        if (pair_less(ov, oj, bv, bj)) {
          bv = ov;
          bj = oj;
        }
      }
      if (l < 16) S.j0[qi] = bj;
    }
    __builtin_amdgcn_wave_barrier();
    // Lower bound of every list over the wave's queries (lane: list j0 + l/16, query l%16).
    for (uint32_t jb = 0; jb < a.Rpad; jb += 4) {
      const uint32_t j = jb + l / 16, q = l % 16;
      float b = INFINITY;
      if (j < R && q < nq) {
        const uint32_t jq = S.j0[q];
        const double Dqj = (double)dq[size_t(q) * a.Rmax + j];
        const double Dq0 = (double)dq[size_t(q) * a.Rmax + jq];
        const double lb = wave_list_bound(Dqj, Dq0, (double)S.eq[q], (double)I.rr[size_t(jq) * R + j],
                                          (double)I.radius[j], a.e, j == jq);
        b = (float)lb;
        if ((double)b > lb) b = nextafterf(b, -INFINITY);
      }
#pragma unroll
      for (int s = 1; s < 16; s <<= 1) b = fminf(b, xor_lane(b, s));
      if ((l % 16) == 0 && j < a.Rpad) {
        lbs[j] = j < R ? b : INFINITY;
        ord[j] = j;
      }
    }
    __builtin_amdgcn_wave_barrier();
    // Bitonic sort of (bound, list) over Rpad entries by the wave.
    for (uint32_t kk = 2; kk <= a.Rpad; kk <<= 1)
      for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
        for (uint32_t i = l; i < a.Rpad; i += kWave) {
          const uint32_t p = i ^ jj;
          if (p > i) {
            const bool up = (i & kk) == 0;
            const float vi = lbs[i], vp = lbs[p];
            const uint32_t oi = ord[i], op = ord[p];
            const bool sw = up ? (vp < vi || (vp == vi && op < oi)) : (vi < vp || (vi == vp && oi < op));
            if (sw) {
              lbs[i] = vp;
              lbs[p] = vi;
              ord[i] = op;
              ord[p] = oi;
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    // 4. Lists in bound order.
    float gmax = INFINITY;
    unsigned long long pairs = 0, steps = 0;
    for (uint32_t idx = 0; idx < R; ++idx) {
      if (lbs[idx] > gmax) break;
      const uint32_t j = ord[idx];
      // Does some query still need list j?
      bool need = false;
      if (l < (int)nq) {
        const uint32_t jq = S.j0[l];
        const double lb = wave_list_bound((double)dq[size_t(l) * a.Rmax + j], (double)dq[size_t(l) * a.Rmax + jq],
                                          (double)S.eq[l], (double)I.rr[size_t(jq) * R + j], (double)I.radius[j], a.e,
                                          j == jq);
        need = lb <= (double)S.gam[l];
      }
      if (__ballot(need) == 0) continue;
      const uint32_t off = I.list_off[j], len = I.list_off[j + 1] - off;
      pairs += (unsigned long long)nq * len;
      steps += (len + kTP - 1) / kTP;
      for (uint32_t p0 = 0; p0 < len; p0 += kTP) {
        f32x4 acc[4];
        uint32_t row[4], pid[4];
        float pxn[4];
        bool live[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
          const uint32_t pp = p0 + 16 * t + (l % 16);
          live[t] = pp < len;
          row[t] = off + (live[t] ? pp : 0);
          pid[t] = I.perm[row[t]];
          pxn[t] = I.xn2[row[t]];
        }
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          if (ks >= KS) break;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            float v[8];
            load8(I.xp + size_t(row[t]) * I.ld, ks * 32 + koff, I.d, live[t], v);
            bf16x8 bh, bl;
            split8(v, bh, bl);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Qh[ks], bh, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Qh[ks], bl, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ql[ks], bh, acc[t], 0, 0, 0);
          }
        }
        float thd[4];
        uint32_t thi[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          thd[i] = S.thd[4 * (l / 16) + i];
          thi[i] = S.thi[4 * (l / 16) + i];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (!live[t]) continue;
          const uint32_t id = pid[t];
          const float xn = pxn[t];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const uint32_t q = 4 * (l / 16) + i;
            if (q >= nq) continue;
            const float v = fmaxf(__builtin_fmaf(-2.f, acc[t][i], qn4[i] + xn), 0.f);
            if (pair_less(v, id, thd[i], thi[i])) {
              const uint32_t s = atomicAdd(&S.cnt[q], 1u);
              S.cd[q][s] = v;
              S.ci[q][s] = id;
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
        // One lane per query folds its staged pairs into its sorted K' list.
        if (l < (int)nq) {
          const uint32_t n = S.cnt[l];
          if (n) {
            float kd = S.thd[l];
            uint32_t ki = S.thi[l];
            for (uint32_t c = 0; c < n; ++c) {
              const float cv = S.cd[l][c];
              const uint32_t cid = S.ci[l][c];
              if (!pair_less(cv, cid, kd, ki)) continue;
              int pos = (int)a.Kp - 1;
              while (pos > 0 && pair_less(cv, cid, S.td[l][pos - 1], S.ti[l][pos - 1])) {
                S.td[l][pos] = S.td[l][pos - 1];
                S.ti[l][pos] = S.ti[l][pos - 1];
                --pos;
              }
              S.td[l][pos] = cv;
              S.ti[l][pos] = cid;
              kd = S.td[l][a.Kp - 1];
              ki = S.ti[l][a.Kp - 1];
            }
            S.thd[l] = kd;
            S.thi[l] = ki;
            S.cnt[l] = 0;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      // Refresh the bounds: gamma_q = sqrt(D~_k + E_q) once a query holds k candidates.
      float g = 0.f;
      if (l < (int)nq) {
        const float v = S.td[l][a.k - 1];
        g = INFINITY;
        if (v < FLT_MAX) {
          const double gd = sqrt(((double)v + (double)S.eq[l]) * (1.0 + 1e-12)) * (1.0 + 1e-12);
          g = (float)gd;
          if ((double)g < gd) g = nextafterf(g, INFINITY);
        }
        S.gam[l] = g;
      }
#pragma unroll
      for (int s = 1; s < kWave; s <<= 1) g = fmaxf(g, xor_lane(g, s));
      gmax = g;
      __builtin_amdgcn_wave_barrier();
    }
    // 5. Candidates and their certificate.
    for (uint32_t i = l; i < nq * a.Kp; i += kWave) {
      const uint32_t q = i / a.Kp, jj = i % a.Kp;
      const size_t gq = size_t(I.qbase + T.q0 + q);
      a.cand_i[gq * a.Kp + jj] = S.ti[q][jj];
      a.cand_d[gq * a.Kp + jj] = S.td[q][jj];
      if (jj == 0) {
        const float vk = S.td[q][a.k - 1], vK = S.td[q][a.Kp - 1];
        const double E = (double)S.eq[q];
        const bool ok = vK >= FLT_MAX || (double)vK - E > ((double)vk + E) * (1.0 + 1e-12);
        a.cert[gq] = ok ? 1u : 0u;
      }
    }
    if (l == 0) {
      atomicAdd(&a.stats[4], pairs);
      atomicAdd(&a.stats[5], 2ull * I.d * (pairs + (unsigned long long)nq * R));
      atomicAdd(&a.stats[6], steps);
      atomicMax(&a.stats[7], wall_clock64() - t_begin);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ void reset_counter(uint32_t *c, unsigned long long *stats) {
  c[0] = 0;
  for (int i = 0; i < kBatchStats; ++i) stats[i] = 0;
}

}  // namespace

int exact_query_common(rbc_exact *h, const float *qd, size_t ldq, uint32_t m, uint32_t k, int precision,
                       float *od, uint32_t *oi, uint32_t ldo, hipStream_t st, rbc_query_stats *stats);

}  // namespace rbc

struct rbc_batched {
  std::vector<rbc_exact *> idx;
  int metric = RBC_L2;
  uint32_t Rmax = 0, Rpad = 0, dpad_max = 0;
  hipStream_t stream = nullptr;
  // Per-call state, reused while the shapes and buffers stay the same.
  struct Plan {
    std::vector<uint32_t> m;
    std::vector<const void *> qptr, optr;
    uint32_t k = 0, qt = 0;
    hipGraphExec_t exec = nullptr;
  } plan;
  rbc::DevBuf<rbc::BInst> dinst;
  rbc::DevBuf<rbc::BTile> dtiles;
  rbc::DevBuf<uint32_t> counter, cand_i, cert, qinst;
  rbc::DevBuf<float> cand_d, scratch, qstage;
  rbc::DevBuf<unsigned long long> stats;
  uint32_t ntiles = 0, mtot = 0, grid = 0;
  int cus = 0, wave_occ = 0;  // CUs; resident wave-kernel workgroups per CU (0: not used)
  std::map<std::vector<uint32_t>, uint32_t> tuned;  // batch shape -> tile size
  // Per index: ||x||^2 of its permuted rows and the largest ||x|| (the MFMA scan's bound).
  std::vector<rbc::DevBuf<float>> xn2;
  std::vector<float> xmax;
  ~rbc_batched() {
    if (plan.exec) (void)hipGraphExecDestroy(plan.exec);
    for (auto *p : idx) rbc_exact_destroy(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace rbc {

namespace {

int batch_tiles(rbc_batched *b, const std::vector<uint32_t> &m, uint32_t qt) {
  std::vector<BTile> tiles;
  for (uint32_t i = 0; i < m.size(); ++i)
    for (uint32_t q0 = 0; q0 < m[i]; q0 += qt) tiles.push_back(BTile{i, q0, std::min(qt, m[i] - q0)});
  // Largest indexes first: their tiles take longest.
  std::stable_sort(tiles.begin(), tiles.end(), [&](const BTile &x, const BTile &y) {
    return b->idx[x.inst]->n > b->idx[y.inst]->n;
  });
  b->ntiles = (uint32_t)tiles.size();
  RBC_TRY(b->dtiles.ensure(std::max<size_t>(1, tiles.size())));
  if (!tiles.empty())
    RBC_HIP_TRY(hipMemcpy(b->dtiles.get(), tiles.data(), tiles.size() * sizeof(BTile), hipMemcpyHostToDevice));
  return RBC_OK;
}

// The wave-autonomous kernel serves L2 with K' <= 32.
bool wave_path(const rbc_batched *b, uint32_t Kp) { return b->metric == RBC_L2 && Kp <= kWK && b->wave_occ > 0; }

// cfg: workgroups per CU for the wave kernel, else unused (the tiles carry the tile size).
int batch_launch(rbc_batched *b, uint32_t k, uint32_t Kp, hipStream_t st, uint32_t cfg) {
  hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, b->counter.get(), b->stats.get());
  if (wave_path(b, Kp)) {
    WaveArgs wa;
    wa.inst = b->dinst.get();
    wa.tiles = b->dtiles.get();
    wa.ntiles = b->ntiles;
    wa.next = b->counter.get();
    wa.scratch = b->scratch.get();
    wa.Rmax = b->Rmax;
    wa.Rpad = b->Rpad;
    wa.k = k;
    wa.Kp = Kp;
    wa.e = fma_rel_err(b->dpad_max);
    wa.cand_i = b->cand_i.get();
    wa.cand_d = b->cand_d.get();
    wa.cert = b->cert.get();
    wa.stats = b->stats.get();
    const uint32_t bpc = std::min(std::max(cfg, 1u), (uint32_t)b->wave_occ);
    hipLaunchKernelGGL(batched_wave_kernel, dim3(b->cus * bpc), dim3(kNT), 4 * wave_lds_bytes(b->Rpad), st, wa);
    RBC_HIP_TRY(hipGetLastError());
  } else {
  BatchArgs a;
  a.inst = b->dinst.get();
  a.tiles = b->dtiles.get();
  a.ntiles = b->ntiles;
  a.next = b->counter.get();
  a.scratch = b->scratch.get();
  a.Rmax = b->Rmax;
  a.Rpad = b->Rpad;
  a.k = k;
  a.Kp = Kp;
  a.e = fma_rel_err(b->dpad_max);
  a.metric = b->metric;
  a.cand_i = b->cand_i.get();
  a.cand_d = b->cand_d.get();
  a.cert = b->cert.get();
  a.qinst = b->qinst.get();
  a.stats = b->stats.get();
  const size_t lds = batch_lds_bytes(b->dpad_max, b->Rpad);
  if (b->metric == RBC_L1)
    hipLaunchKernelGGL(batched_exact_kernel<1>, dim3(b->grid), dim3(kNT), lds, st, a, b->dpad_max);
  else
    hipLaunchKernelGGL(batched_exact_kernel<0>, dim3(b->grid), dim3(kNT), lds, st, a, b->dpad_max);
  RBC_HIP_TRY(hipGetLastError());
  }
  BSelArgs s{b->dinst.get(), b->qinst.get(), b->mtot, k, Kp, 64, b->cand_i.get(), b->cert.get(), b->stats.get()};
  dim3 g(cdiv(b->mtot, 4));
  for (uint32_t P : {64u, 128u}) {
    s.P = P;
    if (b->metric == RBC_L1) {
      if (P == 64) hipLaunchKernelGGL((batched_select<1, 1>), g, dim3(256), 0, st, s);
      else hipLaunchKernelGGL((batched_select<1, 2>), g, dim3(256), 0, st, s);
    } else {
      if (P == 64) hipLaunchKernelGGL((batched_select<0, 1>), g, dim3(256), 0, st, s);
      else hipLaunchKernelGGL((batched_select<0, 2>), g, dim3(256), 0, st, s);
    }
    RBC_HIP_TRY(hipGetLastError());
  }
  return RBC_OK;
}

}  // namespace

}  // namespace rbc

using namespace rbc;

extern "C" {

int rbc_batched_build(uint32_t count, const float *const *x, const uint32_t *n, const uint32_t *d,
                      const uint32_t *ldx, const uint32_t *num_reps, int metric, uint64_t seed, int mem,
                      rbc_batched **out) {
  RBC_ARG_CHECK(out && x && n && d && ldx && num_reps && count, "bad argument");
  *out = nullptr;
  auto b = new rbc_batched();
  b->metric = metric;
  for (uint32_t i = 0; i < count; ++i) {
    if (d[i] > 128) {
      delete b;
      set_error("batched index %u: d=%u above the batched limit of 128", i, d[i]);
      return RBC_ERR_LIMIT;
    }
    rbc_exact *h = nullptr;
    int s = rbc_exact_build(x[i], n[i], d[i], ldx[i], num_reps[i], metric, nullptr, seed + i, mem, &h);
    if (s != RBC_OK) {
      delete b;
      return s;
    }
    b->idx.push_back(h);
    rbc::DevBuf<float> nb;
    if ((s = nb.alloc(h->n)) != RBC_OK) {
      delete b;
      return s;
    }
    hipLaunchKernelGGL(row_norms2, dim3(rbc::cdiv(h->n, 256)), dim3(256), 0, nullptr, h->xp.get(), h->n, h->d, h->ld,
                       nb.get());
    std::vector<float> hn(h->n);
    if (hipMemcpy(hn.data(), nb.get(), size_t(h->n) * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      delete b;
      set_error("copying row norms failed");
      return RBC_ERR_HIP;
    }
    double mx = 0.0;
    for (float v : hn) mx = std::max(mx, (double)v);
    // The stored norms are within 2^-24 (relative) of the FP64 sums.
    const double xm = std::sqrt(mx * (1.0 + 0x1p-23)) * (1.0 + 1e-12);
    float xmf = (float)xm;
    if ((double)xmf < xm) xmf = nextafterf(xmf, FLT_MAX);
    b->xn2.push_back(std::move(nb));
    b->xmax.push_back(xmf);
    b->Rmax = std::max(b->Rmax, h->R);
    b->dpad_max = std::max(b->dpad_max, h->dpad);
  }
  b->Rpad = 1;
  while (b->Rpad < b->Rmax) b->Rpad <<= 1;
  if (b->Rpad > 1024) {
    delete b;
    set_error("batched indexes support at most 1024 representatives (got %u)", b->Rmax);
    return RBC_ERR_LIMIT;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    set_error("device query failed");
    return RBC_ERR_HIP;
  }
  const size_t lds = batch_lds_bytes(b->dpad_max, b->Rpad);
  if ((int)lds > max_lds_bytes()) {
    delete b;
    set_error("batched kernel needs %zu bytes of LDS", lds);
    return RBC_ERR_LIMIT;
  }
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&batched_exact_kernel<0>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes());
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&batched_exact_kernel<1>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes());
  b->grid = (uint32_t)prop.multiProcessorCount;  // one resident workgroup per CU
  b->cus = prop.multiProcessorCount;
  if (metric == RBC_L2) {
    const size_t wlds = 4 * wave_lds_bytes(b->Rpad);
    if ((int)wlds <= max_lds_bytes()) {
      (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&batched_wave_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)wlds);
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batched_wave_kernel, kNT, wlds) == hipSuccess)
        b->wave_occ = std::max(occ, 0);
    }
  }
  const size_t scr = std::max(size_t(b->grid) * kTQ * b->Rmax, size_t(b->cus) * b->wave_occ * 4 * kWQ * b->Rmax);
  int s;
  if ((s = b->scratch.alloc(scr)) || (s = b->counter.alloc(1)) ||
      (s = b->stats.alloc(kBatchStats))) {
    delete b;
    return s;
  }
  *out = b;
  return RBC_OK;
}

int rbc_batched_kquery(rbc_batched *b, const float *const *q, const uint32_t *m, const uint32_t *ldq,
                       uint32_t k, uint32_t *const *nn_ids, float *const *nn_dists, const uint32_t *ldo,
                       int mem, void *stream, rbc_batched_stats *stats) {
  RBC_ARG_CHECK(b && q && m && ldq && nn_ids && nn_dists && ldo, "bad argument");
  RBC_ARG_CHECK(k >= 1 && k <= 56, "k=%u outside [1, 56]", k);
  (void)stream;
  hipStream_t st = b->stream;
  const uint32_t count = (uint32_t)b->idx.size();
  const uint32_t Kp = std::min<uint32_t>(64, k + 8);
  std::vector<uint32_t> mv(m, m + count);
  for (uint32_t i = 0; i < count; ++i) {
    RBC_ARG_CHECK(k <= b->idx[i]->n, "k=%u exceeds index %u's size", k, i);
    RBC_ARG_CHECK(ldq[i] >= b->idx[i]->d && ldo[i] >= k, "bad leading dimension for index %u", i);
  }
  // Stage queries (device memory with ld % 4 == 0 is used in place).
  std::vector<size_t> qoff(count + 1, 0), ooff(count + 1, 0);
  for (uint32_t i = 0; i < count; ++i) {
    qoff[i + 1] = qoff[i] + size_t(m[i]) * pad16(b->idx[i]->d);
    ooff[i + 1] = ooff[i] + size_t(m[i]) * k;
  }
  const bool in_place = mem == RBC_MEM_DEVICE;
  for (uint32_t i = 0; in_place && i < count; ++i)
    RBC_ARG_CHECK(ldq[i] % 4 == 0, "device queries need a leading dimension that is a multiple of 4");
  DevBuf<float> od;
  DevBuf<uint32_t> oi;
  if (!in_place) {
    RBC_TRY(b->qstage.ensure(std::max<size_t>(1, qoff[count])));
    RBC_TRY(od.alloc(std::max<size_t>(1, ooff[count])));
    RBC_TRY(oi.alloc(std::max<size_t>(1, ooff[count])));
    for (uint32_t i = 0; i < count; ++i)
      if (m[i])
        RBC_TRY(rows_in(b->qstage.get() + qoff[i], pad16(b->idx[i]->d) * 4, q[i], size_t(ldq[i]) * 4,
                        size_t(b->idx[i]->d) * 4, m[i], RBC_MEM_HOST, st));
  }
  std::vector<BInst> hi(count);
  uint32_t mtot = 0;
  std::vector<uint32_t> qin;
  for (uint32_t i = 0; i < count; ++i) {
    rbc_exact *h = b->idx[i];
    BInst &I = hi[i];
    I.xp = h->xp.get();
    I.reps = h->reps.get();
    I.list_off = h->list_off.get();
    I.perm = h->perm.get();
    I.inv = h->inv.get();
    I.radius = h->radius.get();
    I.rr = h->rr.get();
    I.xn2 = b->xn2[i].get();
    I.xmax = b->xmax[i];
    I.rhi = h->reps_b.get();
    I.rlo = h->reps_lo.get();
    I.rn2 = h->reps_n2.get();
    I.ldb = h->ldb;
    {
      const double up = 0x1p-23, N = 3.0 * 32.0 * cdiv(h->d, 32) + 1.0;
      I.csplit = 3.1 * 0x1p-16 + 1.02 * N * up / (1.0 - N * up);
    }
    I.n = h->n;
    I.d = h->d;
    I.dpad = h->dpad;
    I.ld = h->ld;
    I.R = h->R;
    I.P = tree_width(h->d);
    I.q = in_place ? q[i] : b->qstage.get() + qoff[i];
    I.ldq = in_place ? ldq[i] : pad16(h->d);
    I.m = m[i];
    I.qbase = mtot;
    I.out_d = in_place ? nn_dists[i] : od.get() + ooff[i];
    I.out_i = in_place ? nn_ids[i] : oi.get() + ooff[i];
    I.ldo = in_place ? ldo[i] : k;
    mtot += m[i];
    qin.insert(qin.end(), m[i], i);
  }
  if (mtot == 0) return RBC_OK;
  RBC_TRY(b->dinst.ensure(count));
  RBC_TRY(upload_host(b->dinst.get(), hi.data(), count * sizeof(BInst), st));
  b->mtot = mtot;
  RBC_TRY(b->qinst.ensure(mtot));
  RBC_TRY(b->cand_i.ensure(size_t(mtot) * Kp));
  RBC_TRY(b->cand_d.ensure(size_t(mtot) * Kp));
  RBC_TRY(b->cert.ensure(mtot));
  RBC_TRY(upload_host(b->qinst.get(), qin.data(), mtot * 4, st));

  // Autotune the tile size once per batch shape.
  std::vector<uint32_t> key(mv);
  key.push_back(k);
  uint32_t qt = 0;
  double tune_ms[3] = {0, 0, 0};
  auto found = b->tuned.find(key);
  const bool wave = wave_path(b, Kp);
  if (found == b->tuned.end()) {
    // Wave kernel: 16-query tiles, and the workgroups per CU are tuned (a quarter, half and
    // all of what fits); otherwise the tile size of the 64-slot tile kernel.
    uint32_t cand_qt[3] = {16, 32, 64};
    if (wave) {
      const uint32_t o = (uint32_t)b->wave_occ;
      cand_qt[0] = std::max(1u, o / 4);
      cand_qt[1] = std::max(1u, o / 2);
      cand_qt[2] = o;
      RBC_TRY(batch_tiles(b, mv, kWQ));
    }
    float best = 1e30f;
    hipEvent_t e0, e1;
    RBC_HIP_TRY(hipEventCreate(&e0));
    RBC_HIP_TRY(hipEventCreate(&e1));
    for (int c = 0; c < 3; ++c) {
      if (!wave) RBC_TRY(batch_tiles(b, mv, cand_qt[c]));
      RBC_HIP_TRY(hipEventRecord(e0, st));
      RBC_TRY(batch_launch(b, k, Kp, st, cand_qt[c]));
      RBC_HIP_TRY(hipEventRecord(e1, st));
      RBC_HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.f;
      RBC_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      tune_ms[c] = ms;
      if (ms < best) {
        best = ms;
        qt = cand_qt[c];
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tuned[key] = qt;
  } else {
    qt = found->second;
  }
  // Graph: reset counter + persistent kernel + re-rank; replayed while nothing changes.
  std::vector<const void *> qp(q, q + count), op;
  for (uint32_t i = 0; i < count; ++i) op.push_back(in_place ? (const void *)nn_ids[i] : (const void *)oi.get());
  bool replay = b->plan.exec && b->plan.m == mv && b->plan.qptr == qp && b->plan.optr == op && b->plan.k == k &&
                b->plan.qt == qt && in_place;
  if (!replay) {
    if (b->plan.exec) {
      (void)hipGraphExecDestroy(b->plan.exec);
      b->plan.exec = nullptr;
    }
    RBC_TRY(batch_tiles(b, mv, wave ? kWQ : qt));
    hipGraph_t graph;
    RBC_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int s = batch_launch(b, k, Kp, st, qt);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (s != RBC_OK) return s;
    RBC_HIP_TRY(ce);
    RBC_HIP_TRY(hipGraphInstantiate(&b->plan.exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    b->plan.m = mv;
    b->plan.qptr = qp;
    b->plan.optr = op;
    b->plan.k = k;
    b->plan.qt = qt;
  }
  hipEvent_t t0, t1;
  RBC_HIP_TRY(hipEventCreate(&t0));
  RBC_HIP_TRY(hipEventCreate(&t1));
  RBC_HIP_TRY(hipEventRecord(t0, st));
  RBC_HIP_TRY(hipGraphLaunch(b->plan.exec, st));
  RBC_HIP_TRY(hipEventRecord(t1, st));
  RBC_HIP_TRY(hipEventSynchronize(t1));
  float ms = 0.f;
  RBC_HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (!in_place) {
    // The graph wrote to this call's staging buffers: keep no replay across calls.
    (void)hipGraphExecDestroy(b->plan.exec);
    b->plan.exec = nullptr;
  }
  unsigned long long hs[kBatchStats];
  RBC_HIP_TRY(hipMemcpy(hs, b->stats.get(), sizeof(hs), hipMemcpyDeviceToHost));
  if (hs[2]) {
    set_error("%llu queries have more than 64 candidates within the FP32 error bound of each other", hs[2]);
    return RBC_ERR_LIMIT;
  }
  // Queries whose K' is not provably complete: re-run them through the single-index path.
  std::vector<uint32_t> hc(mtot);
  RBC_HIP_TRY(hipMemcpy(hc.data(), b->cert.get(), mtot * 4, hipMemcpyDeviceToHost));
  uint32_t uncert = 0;
  for (uint32_t i = 0; i < count; ++i) {
    std::vector<uint32_t> rows;
    for (uint32_t qi = 0; qi < m[i]; ++qi)
      if (!hc[hi[i].qbase + qi]) rows.push_back(qi);
    if (rows.empty()) continue;
    uncert += (uint32_t)rows.size();
    rbc_exact *h = b->idx[i];
    const uint32_t ldh = pad16(h->d);
    DevBuf<float> qs, fd;
    DevBuf<uint32_t> fi;
    RBC_TRY(qs.alloc(rows.size() * ldh));
    RBC_TRY(fd.alloc(rows.size() * k));
    RBC_TRY(fi.alloc(rows.size() * k));
    for (size_t r = 0; r < rows.size(); ++r)
      RBC_HIP_TRY(hipMemcpyAsync(qs.get() + r * ldh, hi[i].q + size_t(rows[r]) * hi[i].ldq, size_t(h->d) * 4,
                                 hipMemcpyDeviceToDevice, st));
    RBC_TRY(exact_query_common(h, qs.get(), ldh, (uint32_t)rows.size(), k, RBC_PREC_FP32, fd.get(), fi.get(), k, st,
                               nullptr));
    for (size_t r = 0; r < rows.size(); ++r) {
      RBC_HIP_TRY(hipMemcpyAsync(hi[i].out_d + size_t(rows[r]) * hi[i].ldo, fd.get() + r * k, k * 4,
                                 hipMemcpyDeviceToDevice, st));
      RBC_HIP_TRY(hipMemcpyAsync(hi[i].out_i + size_t(rows[r]) * hi[i].ldo, fi.get() + r * k, k * 4,
                                 hipMemcpyDeviceToDevice, st));
    }
    RBC_HIP_TRY(hipStreamSynchronize(st));
  }
  if (!in_place) {
    for (uint32_t i = 0; i < count; ++i)
      if (m[i]) {
        RBC_TRY(rows_out(nn_dists[i], size_t(ldo[i]) * 4, od.get() + ooff[i], size_t(k) * 4, size_t(k) * 4, m[i],
                         RBC_MEM_HOST, st));
        RBC_TRY(rows_out(nn_ids[i], size_t(ldo[i]) * 4, oi.get() + ooff[i], size_t(k) * 4, size_t(k) * 4, m[i],
                         RBC_MEM_HOST, st));
      }
  }
  RBC_HIP_TRY(hipStreamSynchronize(st));
  if (stats) {
    stats->ms_graph = ms;
    stats->tile_queries = wave ? kWQ : qt;
    stats->blocks_per_cu = wave ? std::min(qt, (uint32_t)b->wave_occ) : 1u;
    stats->graph_replayed = replay ? 1u : 0u;
    stats->uncertified = uncert;
    stats->tune_ms[0] = tune_ms[0];
    stats->tune_ms[1] = tune_ms[1];
    stats->tune_ms[2] = tune_ms[2];
    stats->guard_fp64 = (double)hs[0];
    stats->pairs_scan = (double)hs[4];
    stats->flops = (double)hs[5];
    stats->steps = (double)hs[6];
    int rate_khz = 0;
    stats->longest_tile_ms =
        hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0) == hipSuccess && rate_khz > 0
            ? (double)hs[7] / rate_khz
            : -1.0;
  }
  return device_check();
}

int rbc_batched_get_index(const rbc_batched *b, uint32_t i, const rbc_exact **out) {
  RBC_ARG_CHECK(b && out, "NULL argument");
  RBC_ARG_CHECK(i < b->idx.size(), "index %u of %zu", i, b->idx.size());
  *out = b->idx[i];
  return RBC_OK;
}

int rbc_batched_destroy(rbc_batched *b) {
  delete b;
  return RBC_OK;
}

}  // extern "C"
