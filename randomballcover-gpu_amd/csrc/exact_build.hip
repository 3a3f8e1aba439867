// Exact-RBC build: representatives -> owner of every point -> stable counting sort
// into contiguous ownership lists (the rep-to-list build) -> permuted FP32 and BF16
// copies, per-list radii in FP64, rep-to-rep distances.
//
// Reference: the exact-search fragments computeRadii (rbc.cu:250-256, unused) and
// groupPoints (deprecated.cu:395-428, uncompiled; a CPU counting sort by owner); the
// one-shot build's scan/scatter compaction (sKernel.cu:14-153) plays the same role for
// its overlapping lists.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <unordered_set>

#include "exact.hpp"
#include "exact_sketch.hpp"

namespace rbc {

namespace {

__global__ void gather_rows(const float *x, size_t ldx, const uint32_t *ids, uint32_t rows, uint32_t d,
                            float *out, size_t ldo) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t tot = size_t(rows) * ldo;
  for (; i < tot; i += size_t(gridDim.x) * blockDim.x) {
    size_t r = i / ldo, c = i % ldo;
    out[i] = c < d ? x[size_t(ids[r]) * ldx + c] : 0.f;
  }
}

// den = 2 sqrt(rr (1 + e)) rounded up to float (0 where rr <= 0: no plane bound).
__global__ void voronoi_den(const float *rr, size_t n, double e, float *den) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const double v = (double)rr[i];
    if (v <= 0.0) {
      den[i] = 0.f;
      continue;
    }
    const double t = 2.0 * sqrt(v * (1.0 + e)) * (1.0 + 1e-15);
    float f = (float)t;
    if ((double)f < t) f = nextafterf(f, INFINITY);
    den[i] = f;
  }
}

__global__ void invert_perm(const uint32_t *perm, uint32_t n, uint32_t *inv) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[perm[i]] = i;
}

__global__ void list_lengths(const uint32_t *off, uint32_t R, uint32_t *len) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < R) len[j] = off[j + 1] - off[j];
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// One wave per permuted row: copy X[perm[pos]] into xp (zero padded), its BF16 copy,
// the FP64 distance to the owning rep (sequential, no contraction), the BF16 squared
// norm and the BF16 quantisation error ||x - bf16(x)||.
__global__ __launch_bounds__(256) void permute_rows(const float *x, size_t ldx, const uint32_t *perm,
                                                    const uint32_t *owner, const float *reps, uint32_t n,
                                                    uint32_t d, uint32_t ld, uint32_t ldb, int metric,
                                                    float *xp, uint16_t *xb, double *own64, float *bnorm,
                                                    float *qerr) {
  uint32_t pos = blockIdx.x * 4 + threadIdx.x / kWave;
  if (pos >= n) return;
  const int l = lane_id();
  uint32_t id = perm[pos];
  const float *src = x + size_t(id) * ldx;
  const float *rp = reps + size_t(owner[id]) * ld;
  double nb = 0.0, qe = 0.0;
  for (uint32_t c = l; c < ldb || c < ld; c += kWave) {
    float v = c < d ? src[c] : 0.f;
    if (c < ld) xp[size_t(pos) * ld + c] = v;
    if (c >= ldb) continue;
    uint16_t h = f32_to_bf16_rne(v);
    xb[size_t(pos) * ldb + c] = h;
    double hv = bf16_to_f32(h);
    nb += hv * hv;
    double e = (double)v - hv;
    qe += e * e;
  }
  for (int s = kWave / 2; s > 0; s >>= 1) {
    nb += __shfl_xor(nb, s, kWave);
    qe += __shfl_xor(qe, s, kWave);
  }
  if (l == 0) {
    double acc = 0.0;
    for (uint32_t c = 0; c < d; ++c) {
      double t = __dsub_rn((double)src[c], (double)rp[c]);
      acc = __dadd_rn(acc, metric == 0 ? __dmul_rn(t, t) : fabs(t));
    }
    own64[pos] = metric == 0 ? sqrt(acc) : acc;
    bnorm[pos] = (float)nb;
    float qf = (float)sqrt(qe);
    if ((double)qf < sqrt(qe)) qf = nextafterf(qf, FLT_MAX);
    qerr[pos] = qf;
  }
}

// One block per list: radius = max FP64 distance (rounded up to float), the largest
// BF16 quantisation error and the largest BF16 squared norm.
__global__ __launch_bounds__(256) void list_stats(const uint32_t *off, const double *own64,
                                                  const float *qerr, const float *bnorm, float *radius,
                                                  float *lqerr, float *lnmax) {
  __shared__ double sm[256];
  __shared__ float sq[256], sn[256];
  uint32_t j = blockIdx.x;
  double m = 0.0;
  float mq = 0.f, mn = 0.f;
  for (uint32_t p = off[j] + threadIdx.x; p < off[j + 1]; p += 256) {
    m = fmax(m, own64[p]);
    mq = fmaxf(mq, qerr[p]);
    mn = fmaxf(mn, bnorm[p]);
  }
  sm[threadIdx.x] = m;
  sq[threadIdx.x] = mq;
  sn[threadIdx.x] = mn;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
      sq[threadIdx.x] = fmaxf(sq[threadIdx.x], sq[threadIdx.x + s]);
      sn[threadIdx.x] = fmaxf(sn[threadIdx.x], sn[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float r = (float)sm[0];
    if ((double)r < sm[0]) r = nextafterf(r, FLT_MAX);
    radius[j] = r;
    lqerr[j] = sq[0];
    lnmax[j] = nextafterf(sn[0], FLT_MAX);
  }
}

// BF16 copies of the reps with their squared norms and quantisation errors.
__global__ void reps_bf16(const float *reps, uint32_t R, uint32_t d, uint32_t ld, uint32_t ldb,
                          uint16_t *rb, float *rbn, float *rqe) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  double nb = 0.0, qe = 0.0;
  for (uint32_t c = 0; c < ldb; ++c) {
    float v = c < d ? reps[size_t(j) * ld + c] : 0.f;
    uint16_t h = f32_to_bf16_rne(v);
    rb[size_t(j) * ldb + c] = h;
    double hv = bf16_to_f32(h);
    nb += hv * hv;
    double e = (double)v - hv;
    qe += e * e;
  }
  rbn[j] = (float)nb;
  float qf = (float)sqrt(qe);
  if ((double)qf < sqrt(qe)) qf = nextafterf(qf, FLT_MAX);
  rqe[j] = qf;
}

// Second BF16 terms of the reps (lo = bf16(r - hi), zero padded to ldb) and ||r||^2 (FP64
// sum, rounded to nearest).
__global__ void reps_split(const float *reps, const uint16_t *rb, uint32_t R, uint32_t d, uint32_t ld,
                           uint32_t ldb, uint16_t *rlo, float *rn2) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  double n2 = 0.0;
  for (uint32_t c = 0; c < ldb; ++c) {
    const float v = c < d ? reps[size_t(j) * ld + c] : 0.f;
    const float hi = bf16_to_f32(rb[size_t(j) * ldb + c]);
    rlo[size_t(j) * ldb + c] = f32_to_bf16_rne(v - hi);  // v - hi is exact in FP32
    n2 += (double)v * (double)v;
  }
  rn2[j] = (float)n2;
}

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

int rep_splits(rbc_exact *h, hipStream_t st) {
  const uint32_t R = h->R;
  RBC_TRY(h->reps_lo.alloc(size_t(R) * h->ldb));
  RBC_TRY(h->reps_n2.alloc(R));
  hipLaunchKernelGGL(reps_split, dim3(cdiv(R, 64)), dim3(64), 0, st, h->reps.get(), h->reps_b.get(), R, h->d,
                     h->ld, h->ldb, h->reps_lo.get(), h->reps_n2.get());
  RBC_HIP_TRY(hipGetLastError());
  std::vector<float> n2(R);
  RBC_HIP_TRY(hipMemcpyAsync(n2.data(), h->reps_n2.get(), size_t(R) * 4, hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  double mx = 0.0;
  for (float v : n2) mx = std::max(mx, (double)v);
  // The stored ||r||^2 is within 2^-24 (relative) of the FP64 sum.
  h->rep_nmax = std::sqrt(mx * (1.0 + 0x1p-23)) * (1.0 + 1e-12);
  return RBC_OK;
}

int voronoi_denominators(rbc_exact *h, hipStream_t st) {
  const size_t n = size_t(h->R) * h->R;
  RBC_TRY(h->rr_den.alloc(n));
  const uint32_t g = (uint32_t)std::min<size_t>(cdivz(n, 256), 4096);
  hipLaunchKernelGGL(voronoi_den, dim3(g), dim3(256), 0, st, h->rr.get(), n, fma_rel_err(h->dpad), h->rr_den.get());
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

// Seeded partial Fisher-Yates: l distinct indices in [0, n) (the role of subRandPerm,
// utils.cu:17-37, with a reproducible generator in place of the clock-seeded rand()).
std::vector<uint32_t> sample_distinct(uint32_t l, uint32_t n, uint64_t seed) {
  std::vector<uint32_t> y(n);
  std::iota(y.begin(), y.end(), 0u);
  uint64_t s = seed ^ 0x5DEECE66Dull;
  for (uint32_t i = 0; i < l && i + 1 < n; ++i) {
    uint64_t r = splitmix64(s);
    uint32_t j = i + uint32_t(r % uint64_t(n - i));
    std::swap(y[i], y[j]);
  }
  y.resize(l);
  return y;
}

int resolve_reps(const uint32_t *rep_ids, uint32_t R, uint32_t n, uint64_t seed,
                 std::vector<uint32_t> &out) {
  if (rep_ids) {
    out.assign(rep_ids, rep_ids + R);
    std::unordered_set<uint32_t> seen;
    for (uint32_t v : out) {
      RBC_ARG_CHECK(v < n, "rep id %u >= n=%u", v, n);
      RBC_ARG_CHECK(seen.insert(v).second, "rep id %u repeated", v);
    }
  } else {
    out = sample_distinct(R, n, seed);
  }
  return RBC_OK;
}

}  // namespace rbc

using namespace rbc;

extern "C" {

int rbc_exact_build(const float *x, uint32_t n, uint32_t d, uint32_t ldx, uint32_t num_reps, int metric,
                    const uint32_t *rep_ids, uint64_t seed, int mem, rbc_exact **out) {
  RBC_ARG_CHECK(out, "NULL output handle");
  *out = nullptr;
  RBC_ARG_CHECK(d >= 1 && d <= RBC_MAX_DIM, "dimension %u outside [1, %d]", d, RBC_MAX_DIM);
  RBC_ARG_CHECK(metric == RBC_L2 || metric == RBC_L1, "unknown metric %d", metric);
  RBC_ARG_CHECK(n >= 1, "empty database");
  RBC_ARG_CHECK(num_reps >= 1 && num_reps <= n, "num_reps=%u outside [1, n=%u]", num_reps, n);
  RBC_ARG_CHECK(num_reps <= 32768, "num_reps=%u above the supported 32768", num_reps);
  hipStream_t st = nullptr;
  auto h = new rbc_exact();
  struct Guard {
    rbc_exact *h;
    ~Guard() { delete h; }
  } guard{h};
  h->n = n;
  h->d = d;
  h->dpad = (d + 3) & ~3u;
  h->ld = pad16(d);
  h->ldb = (d + 31) & ~31u;
  h->R = num_reps;
  h->metric = metric;
  const uint32_t R = num_reps, ld = h->ld;

  DevMat X;
  RBC_TRY(X.stage(x, n, d, ldx, mem, st));
  std::vector<uint32_t> reps;
  RBC_TRY(resolve_reps(rep_ids, R, n, seed, reps));
  RBC_TRY(h->rep_ids.alloc(R));
  RBC_HIP_TRY(hipMemcpy(h->rep_ids.get(), reps.data(), R * 4, hipMemcpyHostToDevice));
  RBC_TRY(h->reps.alloc(size_t(R) * ld));
  hipLaunchKernelGGL(gather_rows, dim3(std::min<size_t>(cdivz(size_t(R) * ld, 256), 8192)), dim3(256), 0,
                     st, X.p, X.ld, h->rep_ids.get(), R, d, h->reps.get(), (size_t)ld);
  RBC_HIP_TRY(hipGetLastError());

  // Owner of every point: nearest rep, FP32 FMA-order distances, lowest rep on ties.
  RBC_TRY(h->owner.alloc(n));
  DevBuf<float> odist;
  RBC_TRY(odist.alloc(n));
  RBC_TRY(nn_device(h->reps.get(), ld, R, nullptr, X.p, X.ld, n, d, metric, odist.get(),
                    h->owner.get(), st));

  // Stable counting sort by owner -> lists.
  RBC_TRY(h->perm.alloc(n));
  RBC_TRY(h->list_off.alloc(R + 1));
  RBC_TRY(counting_sort(h->owner.get(), n, R, h->perm.get(), h->list_off.get(), h->ws, 0, st));
  RBC_TRY(h->inv.alloc(n));
  hipLaunchKernelGGL(invert_perm, dim3(cdiv(n, 256)), dim3(256), 0, st, h->perm.get(), n, h->inv.get());
  RBC_TRY(h->list_len.alloc(R));
  hipLaunchKernelGGL(list_lengths, dim3(cdiv(R, 256)), dim3(256), 0, st, h->list_off.get(), R,
                     h->list_len.get());

  // Permuted FP32 + BF16 copies and per-row FP64 data.
  RBC_TRY(h->xp.alloc(size_t(n) * ld));
  RBC_TRY(h->xb.alloc(size_t(n) * h->ldb));
  RBC_TRY(h->xb_norm.alloc(n));
  DevBuf<double> own64;
  DevBuf<float> qerr;
  RBC_TRY(own64.alloc(n));
  RBC_TRY(qerr.alloc(n));
  hipLaunchKernelGGL(permute_rows, dim3(cdiv(n, 4)), dim3(256), 0, st, X.p, X.ld, h->perm.get(),
                     h->owner.get(), h->reps.get(), n, d, ld, h->ldb, metric, h->xp.get(), h->xb.get(),
                     own64.get(), h->xb_norm.get(), qerr.get());
  RBC_HIP_TRY(hipGetLastError());
  RBC_TRY(h->radius.alloc(R));
  RBC_TRY(h->list_qerr.alloc(R));
  RBC_TRY(h->list_nmax.alloc(R));
  hipLaunchKernelGGL(list_stats, dim3(R), dim3(256), 0, st, h->list_off.get(), own64.get(), qerr.get(),
                     h->xb_norm.get(), h->radius.get(), h->list_qerr.get(), h->list_nmax.get());
  RBC_HIP_TRY(hipGetLastError());
  RBC_TRY(h->reps_b.alloc(size_t(R) * h->ldb));
  RBC_TRY(h->reps_bnorm.alloc(R));
  RBC_TRY(h->reps_qerr.alloc(R));
  hipLaunchKernelGGL(reps_bf16, dim3(cdiv(R, 64)), dim3(64), 0, st, h->reps.get(), R, d, ld, h->ldb,
                     h->reps_b.get(), h->reps_bnorm.get(), h->reps_qerr.get());
  RBC_HIP_TRY(hipGetLastError());

  // Rep-to-rep FP32 distances (for the Voronoi-plane bound).
  RBC_TRY(h->rr.alloc(size_t(R) * R));
  {
    std::vector<ScanItem> iv;
    for (uint32_t q0 = 0; q0 < R; q0 += kTQ)
      iv.push_back(ScanItem{0, R, q0, std::min((uint32_t)kTQ, R - q0), 0});
    DevBuf<ScanItem> items;
    RBC_TRY(upload_items(iv, items, st));
    TileSrc src{h->reps.get(), ld, nullptr, nullptr, h->reps.get(), ld, nullptr, d, h->dpad};
    EpiDense epi{h->rr.get(), R};
    RBC_TRY(launch_tiles_m(metric, src, items.get(), (uint32_t)iv.size(), epi, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
  }
  RBC_TRY(voronoi_denominators(h, st));
  RBC_TRY(rep_splits(h, st));
  RBC_TRY(sketch_build(h, st));
  h->h_list_off.resize(R + 1);
  RBC_HIP_TRY(hipMemcpy(h->h_list_off.data(), h->list_off.get(), (R + 1) * 4, hipMemcpyDeviceToHost));
  RBC_HIP_TRY(hipDeviceSynchronize());
  guard.h = nullptr;
  *out = h;
  return RBC_OK;
}

int rbc_exact_get_info(const rbc_exact *h, rbc_exact_info *info) {
  RBC_ARG_CHECK(h && info, "NULL argument");
  info->n = h->n;
  info->d = h->d;
  info->ld = h->ld;
  info->num_reps = h->R;
  info->metric = (uint32_t)h->metric;
  uint32_t mx = 0, mn = 0xFFFFFFFFu;
  for (uint32_t j = 0; j < h->R; ++j) {
    uint32_t len = h->h_list_off[j + 1] - h->h_list_off[j];
    mx = std::max(mx, len);
    mn = std::min(mn, len);
  }
  info->max_list = mx;
  info->min_list = mn;
  size_t b = 0;
  b += h->xp.n * 4 + h->xb.n * 2 + h->xb_norm.n * 4 + h->perm.n * 4 + h->inv.n * 4 + h->owner.n * 4;
  b += h->reps.n * 4 + h->reps_b.n * 2 + h->rr.n * 4 + h->xs.n * 2 + h->ws.bytes();
  info->device_bytes = b;
  return RBC_OK;
}

int rbc_exact_get_lists(const rbc_exact *h, uint32_t *rep_ids, uint32_t *list_offsets, uint32_t *perm,
                        uint32_t *owner, float *radii) {
  RBC_ARG_CHECK(h, "NULL handle");
  if (rep_ids) RBC_HIP_TRY(hipMemcpy(rep_ids, h->rep_ids.get(), h->R * 4, hipMemcpyDeviceToHost));
  if (list_offsets)
    RBC_HIP_TRY(hipMemcpy(list_offsets, h->list_off.get(), (h->R + 1) * 4, hipMemcpyDeviceToHost));
  if (perm) RBC_HIP_TRY(hipMemcpy(perm, h->perm.get(), size_t(h->n) * 4, hipMemcpyDeviceToHost));
  if (owner) RBC_HIP_TRY(hipMemcpy(owner, h->owner.get(), size_t(h->n) * 4, hipMemcpyDeviceToHost));
  if (radii) RBC_HIP_TRY(hipMemcpy(radii, h->radius.get(), h->R * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

int rbc_exact_destroy(rbc_exact *h) {
  delete h;
  return RBC_OK;
}

}  // extern "C"
