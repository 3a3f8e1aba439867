// FP32 distance tile engine on the vector ALUs.
//
// One workgroup (256 threads, 4 waves) takes a work item = (a run of database points,
// up to 64 queries).  The 64 queries sit in LDS transposed ([dim][query]); the points
// stream through LDS in steps of 64 points x 64 dims.  Each thread owns a 4x4 register
// tile (4 queries x 4 points) and accumulates one fused multiply-add per coordinate in
// coordinate order -- the arithmetic of the reference's kernels (kernels.cu:51-58,
// 183-197, 251-264, 301-310), so its values match the reference's FP32 distances.
// What happens to each 64x64 block of distances is an epilogue policy (top-k, emit
// under a threshold, arg-min, dense store, range count).
#pragma once

#include "rbc_device.hpp"

namespace rbc {

constexpr int kTQ = 64;   // queries per tile
constexpr int kTP = 64;   // points per step
constexpr int kDCH = 64;  // dims per LDS chunk of points
constexpr int kNT = 256;  // threads per workgroup

struct ScanItem {
  uint32_t x_begin;  // first database row (or first entry of xmap)
  uint32_t x_count;  // number of points
  uint32_t q_begin;  // first query row (or first entry of qids)
  uint32_t q_count;  // <= kTQ
  uint32_t aux;      // epilogue-specific tag (segment index, list id, ...)
};

struct TileSrc {
  const float *x;        // database rows
  size_t ldx;            // floats per row (multiple of 4)
  const uint32_t *xmap;  // gathered points: row = xmap[x_begin + p]; else x_begin + p
  const uint32_t *xid;   // id reported for row r: xid[r]; NULL = the row (or xmap) value
  const float *q;        // query rows
  size_t ldq;            // floats per row (multiple of 4)
  const uint32_t *qids;  // gathered queries: row = qids[q_begin + i]; else q_begin + i
  uint32_t d;            // dimension
  uint32_t dpad;         // d rounded up to a multiple of 4
};

// LDS carve-up shared by all epilogues: [qs: dpad x kTQ][xs: kDCH x kTP][qid: kTQ]
// [pid: kTP][epilogue ...].
__host__ __device__ inline size_t tile_lds_base_bytes(uint32_t dpad) {
  return (size_t(dpad) * kTQ + size_t(kDCH) * kTP) * sizeof(float) + (kTQ + kTP) * sizeof(uint32_t);
}

struct TileCtx {
  const ScanItem *it;
  uint32_t *qid_s;  // query row of tile slot i (kDummy if empty)
  uint32_t *pid_s;  // point id of step slot p (kDummy if past the end)
  unsigned char *epi;
  int qg, pg;  // this thread's 4 queries: 4*qg.., 4 points: 4*pg..
};

// Sets up the LDS carve-up of a tile and loads its queries (transposed).
__device__ inline TileCtx tile_setup(const TileSrc &src, const ScanItem &it, float *smem) {
  const int t = threadIdx.x;
  float *qs = smem;
  float *xs = qs + size_t(src.dpad) * kTQ;
  uint32_t *qid_s = reinterpret_cast<uint32_t *>(xs + kDCH * kTP);
  uint32_t *pid_s = qid_s + kTQ;
  TileCtx ctx;
  ctx.it = &it;
  ctx.qid_s = qid_s;
  ctx.pid_s = pid_s;
  ctx.epi = reinterpret_cast<unsigned char *>(pid_s + kTP);
  ctx.qg = t & 15;
  ctx.pg = t >> 4;
  __syncthreads();  // the previous user of this LDS is done
  if (t < kTQ) {
    uint32_t row = kDummy;
    if ((uint32_t)t < it.q_count) row = src.qids ? src.qids[it.q_begin + t] : it.q_begin + t;
    qid_s[t] = row;
  }
  __syncthreads();
  // Queries -> LDS, transposed.  Coordinates past d are stored as 0 and add exactly 0.
  const uint32_t nq4 = src.dpad >> 2;
  for (uint32_t idx = t; idx < nq4 * kTQ; idx += kNT) {
    uint32_t i = idx % kTQ, c4 = idx / kTQ;
    uint32_t row = qid_s[i];
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (row != kDummy) v = *reinterpret_cast<const float4 *>(src.q + size_t(row) * src.ldq + 4 * c4);
    uint32_t c = 4 * c4;
    qs[(c + 0) * kTQ + i] = c + 0 < src.d ? v.x : 0.f;
    qs[(c + 1) * kTQ + i] = c + 1 < src.d ? v.y : 0.f;
    qs[(c + 2) * kTQ + i] = c + 2 < src.d ? v.z : 0.f;
    qs[(c + 3) * kTQ + i] = c + 3 < src.d ? v.w : 0.f;
  }
  return ctx;
}

// Streams points [x_begin, x_begin + x_count) of src past the loaded queries, handing
// every 64-point block of distances to epi.step.
template <int METRIC, class Epi>
__device__ void tile_steps(const TileSrc &src, uint32_t x_begin, uint32_t x_count, Epi &epi, TileCtx &ctx,
                           float *smem) {
  const int t = threadIdx.x;
  float *qs = smem;
  float *xs = qs + size_t(src.dpad) * kTQ;
  uint32_t *pid_s = ctx.pid_s;
  for (uint32_t p0 = 0; p0 < x_count; p0 += kTP) {
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;

    for (uint32_t c0 = 0; c0 < src.dpad; c0 += kDCH) {
      const uint32_t cn = min((uint32_t)kDCH, src.dpad - c0);
      __syncthreads();
      if (c0 == 0 && t < kTP) {
        uint32_t p = p0 + t;
        uint32_t id = kDummy;
        if (p < x_count) {
          uint32_t row = src.xmap ? src.xmap[x_begin + p] : x_begin + p;
          id = src.xid ? src.xid[row] : row;
        }
        pid_s[t] = id;
      }
      for (uint32_t idx = t; idx < (cn >> 2) * kTP; idx += kNT) {
        uint32_t p = idx % kTP, c4 = idx / kTP;
        uint32_t c = c0 + 4 * c4;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (p0 + p < x_count) {
          uint32_t row = src.xmap ? src.xmap[x_begin + p0 + p] : x_begin + p0 + p;
          v = *reinterpret_cast<const float4 *>(src.x + size_t(row) * src.ldx + c);
        }
        xs[(4 * c4 + 0) * kTP + p] = c + 0 < src.d ? v.x : 0.f;
        xs[(4 * c4 + 1) * kTP + p] = c + 1 < src.d ? v.y : 0.f;
        xs[(4 * c4 + 2) * kTP + p] = c + 2 < src.d ? v.z : 0.f;
        xs[(4 * c4 + 3) * kTP + p] = c + 3 < src.d ? v.w : 0.f;
      }
      __syncthreads();
      const float *qrow = qs + size_t(c0) * kTQ + 4 * ctx.qg;
      const float *xrow = xs + 4 * ctx.pg;
      for (uint32_t c = 0; c < cn; ++c) {
        float4 qv = *reinterpret_cast<const float4 *>(qrow + c * kTQ);
        float4 xv = *reinterpret_cast<const float4 *>(xrow + c * kTP);
        float qa[4] = {qv.x, qv.y, qv.z, qv.w};
        float xa[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] = fma_term<METRIC>(xa[b], qa[a], acc[a][b]);
      }
    }
    epi.step(ctx, acc, p0);
  }
}

template <int METRIC, class Epi>
__device__ void tile_run(const TileSrc &src, const ScanItem &it, Epi &epi, float *smem) {
  TileCtx ctx = tile_setup(src, it, smem);
  epi.begin(ctx);
  tile_steps<METRIC>(src, it.x_begin, it.x_count, epi, ctx, smem);
  epi.end(ctx);
}

// ---------------------------------------------------------------- epilogues ----

// Keeps the K (<= 64) smallest (distance, id) pairs of every query of the tile in LDS;
// candidates under the running threshold are staged per step and folded in by one
// thread per query (insertion into the sorted list).  At the end the lists are written to out[(aux*slot_stride + qrow)*ldo].
struct EpiTopK {
  uint32_t K;
  float *out_d;
  uint32_t *out_i;
  uint32_t ldo;
  size_t slot_stride;  // row offset per item.aux (segments); 0 for a single segment

  static __host__ __device__ size_t lds_bytes() {
    return size_t(kTQ) * 64 * 8 * 2 + kTQ * 4 * 3;
  }
  struct L {
    float *td;
    uint32_t *ti;
    float *cd;
    uint32_t *ci;
    uint32_t *cnt;
    float *thd;
    uint32_t *thi;
  };
  __device__ L lay(const TileCtx &c) const {
    L l;
    l.td = reinterpret_cast<float *>(c.epi);
    l.ti = reinterpret_cast<uint32_t *>(l.td + kTQ * 64);
    l.cd = reinterpret_cast<float *>(l.ti + kTQ * 64);
    l.ci = reinterpret_cast<uint32_t *>(l.cd + kTQ * 64);
    l.cnt = l.ci + kTQ * 64;
    l.thd = reinterpret_cast<float *>(l.cnt + kTQ);
    l.thi = reinterpret_cast<uint32_t *>(l.thd + kTQ);
    return l;
  }
  __device__ void begin(const TileCtx &c) {
    L l = lay(c);
    for (int i = threadIdx.x; i < kTQ * 64; i += kNT) {
      l.td[i] = FLT_MAX;
      l.ti[i] = kDummy;
    }
    if (threadIdx.x < kTQ) {
      l.cnt[threadIdx.x] = 0;
      l.thd[threadIdx.x] = FLT_MAX;
      l.thi[threadIdx.x] = kDummy;
    }
  }
  __device__ void step(const TileCtx &c, float (&acc)[4][4], uint32_t) {
    L l = lay(c);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      int qi = 4 * c.qg + a;
      float thd = l.thd[qi];
      uint32_t thi = l.thi[qi];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        uint32_t id = c.pid_s[4 * c.pg + b];
        if (id != kDummy && pair_less(acc[a][b], id, thd, thi)) {
          uint32_t s = atomicAdd(&l.cnt[qi], 1u);
          l.cd[qi * 64 + s] = acc[a][b];
          l.ci[qi * 64 + s] = id;
        }
      }
    }
    __syncthreads();
    // One thread per query folds its staged candidates into the sorted list by insertion.
    if (threadIdx.x < kTQ) {
      const int qi = threadIdx.x;
      const uint32_t n = l.cnt[qi];
      if (n) {
        float kd = l.thd[qi];
        uint32_t ki = l.thi[qi];
        float *td = l.td + qi * 64;
        uint32_t *ti = l.ti + qi * 64;
        for (uint32_t c = 0; c < n; ++c) {
          const float cv = l.cd[qi * 64 + c];
          const uint32_t cid = l.ci[qi * 64 + c];
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
        l.thd[qi] = kd;
        l.thi[qi] = ki;
        l.cnt[qi] = 0;
      }
    }
    __syncthreads();
  }
  __device__ void end(const TileCtx &c) {
    L l = lay(c);
    for (int i = threadIdx.x; i < kTQ * 64; i += kNT) {
      int qi = i / 64, j = i % 64;
      uint32_t row = c.qid_s[qi];
      if (row == kDummy || j >= (int)K) continue;
      size_t o = (size_t(c.it->aux) * slot_stride + row) * ldo + j;
      out_d[o] = l.td[i];
      out_i[o] = l.ti[i];
    }
  }
};

// Appends every (distance, id) with distance <= tau[query] to the query's candidate
// buffer (global, capacity cap); the per-query counter keeps counting past cap so
// the host can detect an overflow.
struct EpiEmit {
  const float *tau;
  float *cand_d;
  uint32_t *cand_i;
  uint32_t *cand_n;
  uint32_t cap;
  static __host__ __device__ size_t lds_bytes() { return kTQ * 4; }
  __device__ void begin(const TileCtx &c) {
    float *th = reinterpret_cast<float *>(c.epi);
    if (threadIdx.x < kTQ) {
      uint32_t row = c.qid_s[threadIdx.x];
      th[threadIdx.x] = row == kDummy ? -FLT_MAX : tau[row];
    }
    __syncthreads();
  }
  __device__ void step(const TileCtx &c, float (&acc)[4][4], uint32_t) {
    const float *th = reinterpret_cast<const float *>(c.epi);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      int qi = 4 * c.qg + a;
      float tq = th[qi];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        uint32_t id = c.pid_s[4 * c.pg + b];
        if (id != kDummy && acc[a][b] <= tq) {
          uint32_t row = c.qid_s[qi];
          uint32_t s = atomicAdd(&cand_n[row], 1u);
          if (s < cap) {
            cand_d[size_t(row) * cap + s] = acc[a][b];
            cand_i[size_t(row) * cap + s] = id;
          }
        }
      }
    }
  }
  __device__ void end(const TileCtx &) {}
};

// Per query: the smallest (distance, id) over the item's points.  Written to
// out[aux*slot_stride + qrow].
struct EpiArgMin {
  float *out_d;
  uint32_t *out_i;
  size_t slot_stride;
  static __host__ __device__ size_t lds_bytes() { return kTQ * 16 * 8; }
  float bd[4];
  uint32_t bi[4];
  __device__ void begin(const TileCtx &) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      bd[a] = FLT_MAX;
      bi[a] = kDummy;
    }
  }
  __device__ void step(const TileCtx &c, float (&acc)[4][4], uint32_t) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        uint32_t id = c.pid_s[4 * c.pg + b];
        if (id != kDummy && pair_less(acc[a][b], id, bd[a], bi[a])) {
          bd[a] = acc[a][b];
          bi[a] = id;
        }
      }
  }
  __device__ void end(const TileCtx &c) {
    float *rd = reinterpret_cast<float *>(c.epi);
    uint32_t *ri = reinterpret_cast<uint32_t *>(rd + kTQ * 16);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      rd[(4 * c.qg + a) * 16 + c.pg] = bd[a];
      ri[(4 * c.qg + a) * 16 + c.pg] = bi[a];
    }
    __syncthreads();
    if (threadIdx.x < kTQ) {
      int qi = threadIdx.x;
      float v = rd[qi * 16];
      uint32_t id = ri[qi * 16];
      for (int g = 1; g < 16; ++g)
        if (pair_less(rd[qi * 16 + g], ri[qi * 16 + g], v, id)) {
          v = rd[qi * 16 + g];
          id = ri[qi * 16 + g];
        }
      uint32_t row = c.qid_s[qi];
      if (row != kDummy) {
        size_t o = size_t(c.it->aux) * slot_stride + row;
        out_d[o] = v;
        out_i[o] = id;
      }
    }
  }
};

// Stores every distance: out[qrow * ldo + col], col = item.aux + p (the item's first
// output column plus the point's position in the item).
struct EpiDense {
  float *out;
  size_t ldo;
  static __host__ __device__ size_t lds_bytes() { return 0; }
  __device__ void begin(const TileCtx &) {}
  __device__ void step(const TileCtx &c, float (&acc)[4][4], uint32_t p0) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      uint32_t row = c.qid_s[4 * c.qg + a];
      if (row == kDummy) continue;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        uint32_t p = p0 + 4 * c.pg + b;
        if (p < c.it->x_count) out[size_t(row) * ldo + c.it->aux + p] = acc[a][b];
      }
    }
  }
  __device__ void end(const TileCtx &) {}
};

// Counts the points with distance < ranges[query] (rangeCountKernel, kernels.cu:455)
// and adds the count to counts[query].
struct EpiRangeCount {
  const float *ranges;
  uint32_t *counts;
  static __host__ __device__ size_t lds_bytes() { return kTQ * 4 * 2; }
  uint32_t cnt[4];
  __device__ void begin(const TileCtx &c) {
    float *rg = reinterpret_cast<float *>(c.epi);
    uint32_t *tot = reinterpret_cast<uint32_t *>(rg + kTQ);
    if (threadIdx.x < kTQ) {
      uint32_t row = c.qid_s[threadIdx.x];
      rg[threadIdx.x] = row == kDummy ? -FLT_MAX : ranges[row];
      tot[threadIdx.x] = 0;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) cnt[a] = 0;
    __syncthreads();
  }
  __device__ void step(const TileCtx &c, float (&acc)[4][4], uint32_t) {
    const float *rg = reinterpret_cast<const float *>(c.epi);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        cnt[a] += (c.pid_s[4 * c.pg + b] != kDummy && acc[a][b] < rg[4 * c.qg + a]) ? 1u : 0u;
  }
  __device__ void end(const TileCtx &c) {
    uint32_t *tot = reinterpret_cast<uint32_t *>(c.epi) + kTQ;
#pragma unroll
    for (int a = 0; a < 4; ++a)
      if (cnt[a]) atomicAdd(&tot[4 * c.qg + a], cnt[a]);
    __syncthreads();
    if (threadIdx.x < kTQ) {
      uint32_t row = c.qid_s[threadIdx.x];
      if (row != kDummy && tot[threadIdx.x]) atomicAdd(&counts[row], tot[threadIdx.x]);
    }
  }
};

template <int METRIC, class Epi>
__global__ __launch_bounds__(kNT) void tile_kernel(TileSrc src, const ScanItem *items, Epi epi) {
  extern __shared__ float smem[];
  ScanItem it = items[blockIdx.x];
  tile_run<METRIC>(src, it, epi, smem);
}

}  // namespace rbc
