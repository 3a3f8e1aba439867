// Brute-force search entry points (the reference's brute.cu) on the FP32 tile engine.
//
// Large databases are split into segments so that a launch has enough work items to
// fill the device even for a few queries; per-segment results are merged by one wave
// per query.
#include <algorithm>
#include <mutex>
#include <vector>

#include "launch.hpp"

namespace rbc {

int metric_setting();

namespace {

// Merges nseg sorted lists of K (dist, id) per query into the final K, written with ldo.
__global__ __launch_bounds__(256) void merge_segments(const float *pd, const uint32_t *pi, uint32_t m,
                                                      uint32_t K, uint32_t nseg, float *od,
                                                      uint32_t *oi, uint32_t ldo) {
  const uint32_t q = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (q >= m) return;
  const int l = lane_id();
  float tv = FLT_MAX;
  uint32_t ti = kDummy;
  for (uint32_t s = 0; s < nseg; ++s) {
    size_t base = (size_t(s) * m + q) * K;
    float cv = l < (int)K ? pd[base + l] : FLT_MAX;
    uint32_t ci = l < (int)K ? pi[base + l] : kDummy;
    wave_merge_lower(tv, ti, cv, ci);
  }
  if (l < (int)K) {
    od[size_t(q) * ldo + l] = tv;
    oi[size_t(q) * ldo + l] = ti;
  }
}

__global__ void merge_argmin(const float *pd, const uint32_t *pi, uint32_t m, uint32_t nseg, float *od,
                             uint32_t *oi) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  float v = FLT_MAX;
  uint32_t id = kDummy;
  for (uint32_t s = 0; s < nseg; ++s) {
    float c = pd[size_t(s) * m + q];
    uint32_t ci = pi[size_t(s) * m + q];
    if (pair_less(c, ci, v, id)) {
      v = c;
      id = ci;
    }
  }
  if (od) od[q] = v;
  oi[q] = id;
}

// Segment length so that the launch has at least ~4 work items per CU.
uint32_t seg_len_for(uint32_t m, uint32_t n) {
  uint32_t qtiles = std::max(1u, cdiv(m, kTQ));
  uint32_t want = 2048;
  uint32_t nseg = std::max(1u, cdiv(want, qtiles));
  nseg = std::min(nseg, std::max(1u, cdiv(n, 4 * kTP)));
  uint32_t len = cdiv(n, nseg);
  return std::max((uint32_t)kTP, cdiv(len, kTP) * kTP);
}

int check_common(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m, uint32_t ldq,
                 uint32_t d, int metric) {
  RBC_ARG_CHECK(d >= 1 && d <= RBC_MAX_DIM, "dimension %u outside [1, %d]", d, RBC_MAX_DIM);
  RBC_ARG_CHECK(metric == RBC_L2 || metric == RBC_L1, "unknown metric %d", metric);
  RBC_ARG_CHECK(ldx >= d && ldq >= d, "leading dimension smaller than d");
  RBC_ARG_CHECK((x || n == 0) && (q || m == 0), "NULL matrix");
  return RBC_OK;
}

// Work items of a brute-force launch on the device: item s * qtiles + t covers queries
// [64 t, 64 t + 64) against segment s (the order grid_items produces on the host).
__global__ void grid_items_k(uint32_t m, uint32_t n, uint32_t seg_len, uint32_t nseg, uint32_t qtiles,
                             ScanItem *items) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg * qtiles) return;
  const uint32_t s = i / qtiles, q0 = (i % qtiles) * kTQ;
  ScanItem it;
  it.x_begin = s * seg_len;
  it.x_count = n ? min(seg_len, n - s * seg_len) : 0;
  it.q_begin = q0;
  it.q_count = min((uint32_t)kTQ, m - q0);
  it.aux = s;
  items[i] = it;
}

// ---- allocation-free brute force --------------------------------------------------------
// Scratch of rbc_brute_k_noalloc: the work items and the per-segment partial top-k lists.
// Each device holds its own copy of this array from module load until the process ends.
// Host code never passes the array's name to a kernel (that is a host-side stand-in, not a
// device address): it uses the address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of the current device's scratch array, looked up once per device.
int scratch_address(void **out) {
  int dev = 0;
  RBC_HIP_TRY(hipGetDevice(&dev));
  RBC_ARG_CHECK(dev >= 0 && dev < kMaxDevices, "device %d is beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    RBC_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RBC_ARG_CHECK(p != nullptr, "scratch array has no device address");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return RBC_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

int brute_k_device(const float *x, size_t ldx, uint32_t n, const float *q, size_t ldq, uint32_t m,
                   uint32_t d, int metric, uint32_t K, float *od, uint32_t *oi, uint32_t ldo,
                   hipStream_t st) {
  uint32_t seg = seg_len_for(m, n);
  std::vector<ScanItem> iv = grid_items(m, n, seg);
  uint32_t nseg = n ? cdiv(n, seg) : 1;
  DevBuf<ScanItem> items;
  RBC_TRY(upload_items(iv, items, st));
  DevBuf<float> pd;
  DevBuf<uint32_t> pi;
  RBC_TRY(pd.alloc(size_t(nseg) * m * K));
  RBC_TRY(pi.alloc(size_t(nseg) * m * K));
  TileSrc src{x, ldx, nullptr, nullptr, q, ldq, nullptr, d, (d + 3) & ~3u};
  EpiTopK epi{K, pd.get(), pi.get(), K, m};
  RBC_TRY(launch_tiles_m(metric, src, items.get(), (uint32_t)iv.size(), epi, st));
  hipLaunchKernelGGL(merge_segments, dim3(cdiv(m, 4)), dim3(256), 0, st, pd.get(), pi.get(), m, K,
                     nseg, od, oi, ldo);
  RBC_HIP_TRY(hipGetLastError());
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int merge_slots(const float *pd, const uint32_t *pi, uint32_t m, uint32_t K, uint32_t nslot, float *od,
                uint32_t *oi, uint32_t ldo, hipStream_t st) {
  if (m == 0) return RBC_OK;
  if (K == 1) {
    hipLaunchKernelGGL(merge_argmin, dim3(cdiv(m, 256)), dim3(256), 0, st, pd, pi, m, nslot, od, oi);
  } else {
    hipLaunchKernelGGL(merge_segments, dim3(cdiv(m, 4)), dim3(256), 0, st, pd, pi, m, K, nslot, od, oi, ldo);
  }
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int nn_device(const float *x, size_t ldx, uint32_t n, const uint32_t *xid, const float *q,
              size_t ldq, uint32_t m, uint32_t d, int metric, float *od, uint32_t *oi,
              hipStream_t st) {
  uint32_t seg = seg_len_for(m, n);
  std::vector<ScanItem> iv = grid_items(m, n, seg);
  uint32_t nseg = n ? cdiv(n, seg) : 1;
  DevBuf<ScanItem> items;
  RBC_TRY(upload_items(iv, items, st));
  DevBuf<float> pd;
  DevBuf<uint32_t> pi;
  RBC_TRY(pd.alloc(size_t(nseg) * m));
  RBC_TRY(pi.alloc(size_t(nseg) * m));
  TileSrc src{x, ldx, nullptr, xid, q, ldq, nullptr, d, (d + 3) & ~3u};
  EpiArgMin epi;
  epi.out_d = pd.get();
  epi.out_i = pi.get();
  epi.slot_stride = m;
  RBC_TRY(launch_tiles_m(metric, src, items.get(), (uint32_t)iv.size(), epi, st));
  hipLaunchKernelGGL(merge_argmin, dim3(cdiv(m, 256)), dim3(256), 0, st, pd.get(), pi.get(), m, nseg,
                     od, oi);
  RBC_HIP_TRY(hipGetLastError());
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

}  // namespace rbc

using namespace rbc;

extern "C" {

int rbc_brute_k(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m, uint32_t ldq,
                uint32_t d, int metric, uint32_t k, uint32_t *nn_ids, float *nn_dists, uint32_t ldo,
                int mem, void *stream) {
  RBC_TRY(check_common(x, n, ldx, q, m, ldq, d, metric));
  RBC_ARG_CHECK(k >= 1 && k <= KMAX, "k=%u outside [1, %d]", k, KMAX);
  RBC_ARG_CHECK(ldo >= k, "output leading dimension %u < k=%u", ldo, k);
  RBC_ARG_CHECK(nn_ids && nn_dists, "NULL output");
  if (m == 0) return RBC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevMat X, Q;
  RBC_TRY(X.stage(x, n, d, ldx, mem, st));
  RBC_TRY(Q.stage(q, m, d, ldq, mem, st));
  if (mem == RBC_MEM_DEVICE) {
    return brute_k_device(X.p, X.ld, n, Q.p, Q.ld, m, d, metric, k, nn_dists, nn_ids, ldo, st);
  }
  DevBuf<float> od;
  DevBuf<uint32_t> oi;
  RBC_TRY(od.alloc(size_t(m) * k));
  RBC_TRY(oi.alloc(size_t(m) * k));
  RBC_TRY(brute_k_device(X.p, X.ld, n, Q.p, Q.ld, m, d, metric, k, od.get(), oi.get(), k, st));
  RBC_TRY(rows_out(nn_dists, size_t(ldo) * 4, od.get(), size_t(k) * 4, size_t(k) * 4, m, mem, st));
  RBC_TRY(rows_out(nn_ids, size_t(ldo) * 4, oi.get(), size_t(k) * 4, size_t(k) * 4, m, mem, st));
  return RBC_OK;
}

int rbc_brute_search(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                     uint32_t ldq, uint32_t d, int metric, uint32_t *nn_ids, float *nn_dists,
                     int mem, void *stream) {
  RBC_TRY(check_common(x, n, ldx, q, m, ldq, d, metric));
  RBC_ARG_CHECK(nn_ids, "NULL output");
  if (m == 0) return RBC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevMat X, Q;
  RBC_TRY(X.stage(x, n, d, ldx, mem, st));
  RBC_TRY(Q.stage(q, m, d, ldq, mem, st));
  if (mem == RBC_MEM_DEVICE)
    return nn_device(X.p, X.ld, n, nullptr, Q.p, Q.ld, m, d, metric, nn_dists, nn_ids, st);
  DevBuf<float> od;
  DevBuf<uint32_t> oi;
  RBC_TRY(od.alloc(m));
  RBC_TRY(oi.alloc(m));
  RBC_TRY(nn_device(X.p, X.ld, n, nullptr, Q.p, Q.ld, m, d, metric, od.get(), oi.get(), st));
  RBC_HIP_TRY(hipMemcpy(nn_ids, oi.get(), size_t(m) * 4, hipMemcpyDeviceToHost));
  if (nn_dists) RBC_HIP_TRY(hipMemcpy(nn_dists, od.get(), size_t(m) * 4, hipMemcpyDeviceToHost));
  return RBC_OK;
}

int rbc_brute_range_count(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m,
                          uint32_t ldq, uint32_t d, int metric, const float *ranges, uint32_t *counts,
                          int mem, void *stream) {
  RBC_TRY(check_common(x, n, ldx, q, m, ldq, d, metric));
  RBC_ARG_CHECK(ranges && counts, "NULL ranges or counts");
  if (m == 0) return RBC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevMat X, Q;
  RBC_TRY(X.stage(x, n, d, ldx, mem, st));
  RBC_TRY(Q.stage(q, m, d, ldq, mem, st));
  DevBuf<float> rg;
  DevBuf<uint32_t> cnt;
  RBC_TRY(rg.alloc(m));
  RBC_TRY(cnt.alloc(m));
  RBC_TRY(copy_in(rg.get(), ranges, size_t(m) * 4, mem, st));
  RBC_TRY(fill_u32(cnt.get(), 0, m, st));
  uint32_t seg = seg_len_for(m, n);
  std::vector<ScanItem> iv = grid_items(m, n, seg);
  DevBuf<ScanItem> items;
  RBC_TRY(upload_items(iv, items, st));
  TileSrc src{X.p, X.ld, nullptr, nullptr, Q.p, Q.ld, nullptr, d, (d + 3) & ~3u};
  EpiRangeCount epi;
  epi.ranges = rg.get();
  epi.counts = cnt.get();
  RBC_TRY(launch_tiles_m(metric, src, items.get(), (uint32_t)iv.size(), epi, st));
  RBC_TRY(copy_out(counts, cnt.get(), size_t(m) * 4, mem, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int rbc_brute_k_noalloc(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m, uint32_t ldq,
                        uint32_t d, int metric, uint32_t k, uint32_t *nn_ids, float *nn_dists, uint32_t ldo,
                        void *stream) {
  RBC_TRY(check_common(x, n, ldx, q, m, ldq, d, metric));
  RBC_ARG_CHECK(k >= 1 && k <= KMAX, "k=%u outside [1, %d]", k, KMAX);
  RBC_ARG_CHECK(ldo >= k, "output leading dimension %u < k=%u", ldo, k);
  RBC_ARG_CHECK(nn_ids && nn_dists, "NULL output");
  // The tile engine reads rows with 16-byte loads; the allocating entry point re-pitches
  // other layouts into a copy, which this one cannot make.
  RBC_ARG_CHECK(ldx % 4 == 0 && ldq % 4 == 0, "leading dimensions must be multiples of 4 (ldx=%u, ldq=%u)",
                ldx, ldq);
  if (m == 0) return RBC_OK;
  if (n) RBC_TRY(check_device_ptr(x, "x"));
  RBC_TRY(check_device_ptr(q, "q"));
  RBC_TRY(check_device_ptr(nn_ids, "nn_ids"));
  RBC_TRY(check_device_ptr(nn_dists, "nn_dists"));
  // Segments as the allocating entry point cuts them, fewer if their partial lists would
  // not fit the scratch array.
  const uint32_t qtiles = cdiv(m, kTQ);
  const size_t part = size_t(m) * k * 4;  // one segment's partial distances (ids: the same)
  uint32_t seg = seg_len_for(m, n);
  uint32_t nseg = n ? cdiv(n, seg) : 1;
  auto need = [&](uint32_t ns) { return align256(size_t(ns) * qtiles * sizeof(ScanItem)) + 2 * align256(ns * part); };
  if (need(nseg) > kScratchBytes) {
    const size_t fit = kScratchBytes / (2 * part + 512);
    RBC_ARG_CHECK(fit >= 1 && need(1) <= kScratchBytes,
                  "m=%u queries at k=%u need %zu bytes of scratch; the allocation-free entry point has %zu "
                  "(split the queries)",
                  m, k, need(1), kScratchBytes);
    nseg = (uint32_t)std::min<size_t>(nseg, fit);
    while (nseg > 1 && need(nseg) > kScratchBytes) --nseg;
    seg = std::max((uint32_t)kTP, cdiv(cdiv(n, nseg), kTP) * kTP);
    nseg = n ? cdiv(n, seg) : 1;
  }
  void *base = nullptr;
  RBC_TRY(scratch_address(&base));
  unsigned char *sp = static_cast<unsigned char *>(base);
  ScanItem *items = reinterpret_cast<ScanItem *>(sp);
  float *pd = reinterpret_cast<float *>(sp + align256(size_t(nseg) * qtiles * sizeof(ScanItem)));
  uint32_t *pi = reinterpret_cast<uint32_t *>(reinterpret_cast<unsigned char *>(pd) + align256(nseg * part));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t nitems = nseg * qtiles;
  hipLaunchKernelGGL(grid_items_k, dim3(cdiv(nitems, 256)), dim3(256), 0, st, m, n, seg, nseg, qtiles, items);
  RBC_HIP_TRY(hipGetLastError());
  // Every slot of the partial lists is written by the tile epilogue before the merge reads
  // it, so nothing in the scratch array has to start at zero.
  TileSrc src{x, ldx, nullptr, nullptr, q, ldq, nullptr, d, (d + 3) & ~3u};
  EpiTopK epi{k, pd, pi, k, m};
  RBC_TRY(launch_tiles_m(metric, src, items, nitems, epi, st));
  hipLaunchKernelGGL(merge_segments, dim3(cdiv(m, 4)), dim3(256), 0, st, pd, pi, m, k, nseg, nn_dists, nn_ids,
                     ldo);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

/* Host routine, as in the reference (brute.cu:91-112, utils.cu:183-191). */
int rbc_brute_cpu(const float *x, uint32_t n, uint32_t ldx, const float *q, uint32_t m, uint32_t ldq,
                  uint32_t d, int metric, uint32_t *nn_ids) {
  RBC_TRY(check_common(x, n, ldx, q, m, ldq, d, metric));
  RBC_ARG_CHECK(nn_ids, "NULL output");
  for (uint32_t i = 0; i < m; ++i) {
    float best = FLT_MAX;
    uint32_t arg = 0;
    for (uint32_t j = 0; j < n; ++j) {
      float acc = 0.f;
      for (uint32_t c = 0; c < d; ++c) {
        float t = q[size_t(i) * ldq + c] - x[size_t(j) * ldx + c];
        volatile float term = metric == RBC_L2 ? t * t : fabsf(t);
        acc = acc + term;
      }
      if (acc < best) {
        best = acc;
        arg = j;
      }
    }
    nn_ids[i] = arg;
  }
  return RBC_OK;
}

}  // extern "C"
