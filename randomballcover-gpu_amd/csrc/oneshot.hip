// One-shot RBC with the reference's semantics (rbc.cu): buildRBC / queryRBC / kqueryRBC /
// destroyRBC over the byte-identical rbcStruct.
//
// Build (rbc.cu:126-208): for each chunk of reps, the rep x point distance block (the
// dist1Kernel role, FP32 FMA order), the per-rep bisection range (findRangeKernel,
// kernels.cu:322-414, with its four-row block coupling and an iteration cap the
// reference lacks), then the flag -> scan -> scatter compaction (rangeSearchKernel +
// sumKernel + buildMapKernel + getCountsKernel, kernels.cu:417-423, sKernel.cu:14-153)
// written as a per-row segmented count, one global exclusive scan and a scatter with
// in-block wave scans.  Lists come out in ascending point index, as in the reference.
// Query (rbc.cu:19-123): nearest rep of each query, queries grouped by rep (the
// buildQMap role, here a stable counting sort on the device), then every group scans
// its rep's list through the xMap gather -- or, with a general computation plan, the
// lists the plan names for it.
#include <algorithm>
#include <vector>

#include "exact.hpp"

namespace rbc {

int metric_setting();
std::vector<uint32_t> sample_distinct(uint32_t l, uint32_t n, uint64_t seed);
int resolve_reps(const uint32_t *rep_ids, uint32_t R, uint32_t n, uint64_t seed,
                 std::vector<uint32_t> &out);

namespace {

constexpr uint32_t kMaxBisect = 1000;
constexpr uint32_t kSeg = 2048;  // columns per compaction segment

__device__ float block_min(float v, float *red) {
  for (int s = kWave / 2; s > 0; s >>= 1) v = fminf(v, __shfl_xor(v, s, kWave));
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  float r = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
  __syncthreads();
  return r;
}
__device__ float block_max(float v, float *red) {
  for (int s = kWave / 2; s > 0; s >>= 1) v = fmaxf(v, __shfl_xor(v, s, kWave));
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  float r = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  __syncthreads();
  return r;
}
__device__ uint32_t block_sum(uint32_t v, uint32_t *red) {
  for (int s = kWave / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, kWave);
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  uint32_t r = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return r;
}

// One block (256 threads) per group of four consecutive rows of the distance block D
// (rows x n, leading dimension n).  Rows >= rows_valid are the padding rows.
__global__ __launch_bounds__(256) void find_range(const float *D, uint32_t rows_valid, uint32_t n,
                                                  uint32_t cnt_want, uint32_t max_iters, float *ranges,
                                                  uint32_t *nonconv) {
  __shared__ float redf[4];
  __shared__ uint32_t redu[4];
  const uint32_t r0 = blockIdx.x * 4;
  const uint32_t LB = (90 * cnt_want) / 100, UB = cnt_want;  // kernels.cu:329-330
  float smin[4], smax[4], rg[4] = {0.f, 0.f, 0.f, 0.f};
  int cont[4];
  for (int ro = 0; ro < 4; ++ro) {
    float mn = FLT_MAX, mx = 0.f;  // kernels.cu:335-336
    if (r0 + ro < rows_valid) {
      const float *row = D + size_t(r0 + ro) * n;
      for (uint32_t c = threadIdx.x; c < n; c += 256) {
        float t = row[c];
        mn = t <= mn ? t : mn;
        mx = t > mx ? t : mx;
      }
    }
    smin[ro] = block_min(mn, redf);
    smax[ro] = block_max(mx, redf);
    cont[ro] = 1;
  }
  uint32_t it = 0;
  do {
    ++it;
    for (int ro = 0; ro < 4; ++ro) {
      const bool valid = r0 + ro < rows_valid;
      if (cont[ro]) rg[ro] = (smax[ro] + smin[ro]) / 2.0f;  // kernels.cu:372-373
      uint32_t c = 0;
      if (valid) {
        const float *row = D + size_t(r0 + ro) * n;
        for (uint32_t j = threadIdx.x; j < n; j += 256) c += row[j] <= rg[ro];  // kernels.cu:377
      }
      uint32_t cnt = block_sum(c, redu);
      if (cnt < cnt_want)  // kernels.cu:390-395
        smin[ro] = rg[ro];
      else
        smax[ro] = rg[ro];
      cont[ro] = valid && (cnt < LB || cnt > UB);  // kernels.cu:398-399
    }
    // kernels.cu:403-407: the OR-reduction also rewrites rows 0 and 1's flags.
    cont[0] |= cont[2];
    cont[1] |= cont[3];
    cont[0] |= cont[1];
  } while (cont[0] && it < max_iters);
  if (threadIdx.x == 0) {
    for (int ro = 0; ro < 4; ++ro)
      if (r0 + ro < rows_valid) ranges[r0 + ro] = rg[ro];
    if (cont[0]) atomicAdd(nonconv, 1u);
  }
}

// Per (row, segment): number of columns j < n with D[row][j] < range[row].
__global__ __launch_bounds__(256) void flag_count(const float *D, uint32_t n, const float *ranges,
                                                  uint32_t nseg, uint32_t *cnt) {
  __shared__ uint32_t redu[4];
  const uint32_t row = blockIdx.y, seg = blockIdx.x;
  const float rg = ranges[row];
  uint32_t c = 0;
  for (uint32_t j = seg * kSeg + threadIdx.x; j < min(n, (seg + 1) * kSeg); j += 256)
    c += D[size_t(row) * n + j] < rg;
  uint32_t tot = block_sum(c, redu);
  if (threadIdx.x == 0) cnt[size_t(row) * nseg + seg] = tot;
}

// Scatter: map[row_abs][pos] = j for flagged j, pos = row-local exclusive prefix.
__global__ __launch_bounds__(256) void flag_scatter(const float *D, uint32_t n, const float *ranges,
                                                    uint32_t nseg, const uint32_t *off, uint32_t row0,
                                                    uint32_t *map, uint32_t ldmap) {
  __shared__ uint32_t wsum[4];
  const uint32_t row = blockIdx.y, seg = blockIdx.x;
  const float rg = ranges[row];
  uint32_t base = off[size_t(row) * nseg + seg] - off[size_t(row) * nseg];
  const int w = threadIdx.x / kWave, l = lane_id();
  for (uint32_t j0 = seg * kSeg; j0 < min(n, (seg + 1) * kSeg); j0 += 256) {
    uint32_t j = j0 + threadIdx.x;
    bool f = j < n && j < (seg + 1) * kSeg && D[size_t(row) * n + j] < rg;
    uint64_t b = __ballot(f);
    uint32_t before = (uint32_t)__popcll(b & ((1ull << l) - 1));
    if (l == 0) wsum[w] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t wb = 0, tot = 0;
    for (int i = 0; i < 4; ++i) {
      if (i < w) wb += wsum[i];
      tot += wsum[i];
    }
    uint32_t pos = base + wb + before;
    if (f && pos < ldmap) map[size_t(row0 + row) * ldmap + pos] = j;
    base += tot;
    __syncthreads();
  }
}

__global__ void row_totals(const uint32_t *off, uint32_t rows, uint32_t nseg, uint32_t *tot) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) tot[r] = off[size_t(r + 1) * nseg] - off[size_t(r) * nseg];
}

}  // namespace

// General computation plan (compPlan, defs.h:86-92, built by initCompPlan from a charMatrix,
// rbc.cu:375-431): the queries whose nearest rep is i scan the lists of every rep j with
// plan[i * ldp + j] != 0, in ascending j (planNNKernel / planKNNKernel, kernels.cu:15-161).
// The result is the K best of the union of those lists taken as a multiset: a point held
// by two planned lists is seen twice, as in the reference, which keeps no record of the
// points already merged.  Work items are (group, planned list, <= kTQ queries); the k-th
// planned list of a group writes result slot k.  Slots are merged in passes of kPlanSlots
// - 1 lists, slot 0 carrying the running result, so the scratch stays kPlanSlots x m x K
// for any plan (fullIntersection, rbc.cu:305-312, included).
constexpr uint32_t kPlanSlots = 8;

static int oneshot_query_plan(const rbcStruct *s, const float *Qp, size_t ldQ, uint32_t m, uint32_t K,
                       const uint32_t *qorder, const std::vector<uint32_t> &hqoff, const uint8_t *plan,
                       uint32_t ldp, const std::vector<uint32_t> &hxo, const std::vector<uint32_t> &hxl,
                       int metric, float *od, uint32_t *oi, hipStream_t st) {
  const uint32_t d = s->dx.c, R = s->dr.r;
  // Planned lists of each group that has queries.
  std::vector<std::vector<uint32_t>> lists(R);
  size_t most = 0;
  for (uint32_t i = 0; i < R; ++i) {
    if (hqoff[i + 1] == hqoff[i]) continue;
    for (uint32_t j = 0; j < R; ++j)
      if (plan[size_t(i) * ldp + j]) lists[i].push_back(j);
    most = std::max(most, lists[i].size());
  }
  const uint32_t per = kPlanSlots - 1;
  const size_t slot = size_t(m) * K;
  DevBuf<float> pd;
  DevBuf<uint32_t> pi;
  RBC_TRY(pd.alloc(kPlanSlots * slot));
  RBC_TRY(pi.alloc(kPlanSlots * slot));
  RBC_TRY(fill_f32(pd.get(), FLT_MAX, slot, st));
  RBC_TRY(fill_u32(pi.get(), kDummy, slot, st));
  TileSrc src{s->dx.mat, s->dx.ld, s->dxMap.mat, nullptr, Qp, ldQ, qorder, d, (d + 3) & ~3u};
  DevBuf<ScanItem> items;
  for (size_t first = 0; first < most; first += per) {
    std::vector<ScanItem> iv;
    for (uint32_t i = 0; i < R; ++i) {
      const uint32_t cnt = hqoff[i + 1] - hqoff[i];
      for (size_t k = first; k < std::min(first + per, lists[i].size()); ++k) {
        const uint32_t j = lists[i][k];
        for (uint32_t t = 0; t < cnt; t += kTQ) {
          ScanItem it;
          it.x_begin = hxo[j];
          it.x_count = hxl[j];
          it.q_begin = hqoff[i] + t;
          it.q_count = std::min<uint32_t>(kTQ, cnt - t);
          it.aux = 1 + (uint32_t)(k - first);
          iv.push_back(it);
        }
      }
    }
    RBC_TRY(fill_f32(pd.get() + slot, FLT_MAX, per * slot, st));
    RBC_TRY(fill_u32(pi.get() + slot, kDummy, per * slot, st));
    RBC_TRY(upload_items(iv, items, st));
    if (K == 1) {
      EpiArgMin epi;
      epi.out_d = pd.get();
      epi.out_i = pi.get();
      epi.slot_stride = m;
      RBC_TRY(launch_tiles_m(metric, src, items.get(), (uint32_t)iv.size(), epi, st));
    } else {
      EpiTopK epi{K, pd.get(), pi.get(), K, m};
      RBC_TRY(launch_tiles_m(metric, src, items.get(), (uint32_t)iv.size(), epi, st));
    }
    // In place: each query's slots are read by one wave (one thread for K = 1) before it
    // writes slot 0.
    RBC_TRY(merge_slots(pd.get(), pi.get(), m, K, kPlanSlots, pd.get(), pi.get(), K, st));
    // The items buffer is reused by the next pass.
    RBC_HIP_TRY(hipStreamSynchronize(st));
  }
  RBC_HIP_TRY(hipMemcpyAsync(od, pd.get(), slot * 4, hipMemcpyDeviceToDevice, st));
  RBC_HIP_TRY(hipMemcpyAsync(oi, pi.get(), slot * 4, hipMemcpyDeviceToDevice, st));
  return RBC_OK;
}

int oneshot_query_common(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *s, uint32_t K,
                         uint32_t *ids, float *dists, uint32_t ldo, int mem, hipStream_t st,
                         const uint8_t *plan = nullptr, uint32_t ldp = 0) {
  RBC_ARG_CHECK(s && s->dx.mat && s->dr.mat && s->dxMap.mat && s->groupCount, "RBC not built");
  const uint32_t d = s->dx.c, R = s->dr.r;
  RBC_ARG_CHECK(ldq >= d, "leading dimension %u < dimension %u", ldq, d);
  RBC_ARG_CHECK(ids && (dists || K == 1), "NULL output");
  const int metric = metric_setting();
  DevMat Q;
  RBC_TRY(Q.stage(q, m, d, ldq, mem, st));
  Workspace ws;
  // Nearest rep of every query (computeReps, rbc.cu:233-247).
  DevBuf<uint32_t> near;
  DevBuf<float> nd;
  RBC_TRY(near.alloc(m));
  RBC_TRY(nd.alloc(m));
  RBC_TRY(nn_device(s->dr.mat, s->dr.ld, R, nullptr, Q.p, Q.ld, m, d, metric, nd.get(), near.get(), st));
  // Group queries by rep (computeCounts + buildQMap, rbc.cu:260-294).
  DevBuf<uint32_t> qorder, qoff, xoff, xlen;
  RBC_TRY(qorder.alloc(m));
  RBC_TRY(qoff.alloc(R + 1));
  RBC_TRY(counting_sort(near.get(), m, R, qorder.get(), qoff.get(), ws, 0, st));
  std::vector<uint32_t> hxo(R), hxl(R);
  for (uint32_t r = 0; r < R; ++r) {
    hxo[r] = r * s->dxMap.ld;
    hxl[r] = std::min(s->groupCount[r], s->dxMap.ld);
  }
  DevBuf<float> od;
  DevBuf<uint32_t> oi;
  RBC_TRY(od.alloc(size_t(m) * K));
  RBC_TRY(oi.alloc(size_t(m) * K));
  if (plan) {
    std::vector<uint32_t> hqoff(R + 1);
    RBC_HIP_TRY(hipMemcpyAsync(hqoff.data(), qoff.get(), (R + 1) * 4, hipMemcpyDeviceToHost, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
    RBC_TRY(oneshot_query_plan(s, Q.p, Q.ld, m, K, qorder.get(), hqoff, plan, ldp, hxo, hxl, metric, od.get(),
                               oi.get(), st));
    RBC_TRY(rows_out(ids, size_t(ldo) * 4, oi.get(), size_t(K) * 4, size_t(K) * 4, m, mem, st));
    if (dists) RBC_TRY(rows_out(dists, size_t(ldo) * 4, od.get(), size_t(K) * 4, size_t(K) * 4, m, mem, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
    return RBC_OK;
  }
  RBC_TRY(xoff.alloc(R));
  RBC_TRY(xlen.alloc(R));
  RBC_TRY(upload_host(xoff.get(), hxo.data(), R * 4, st));
  RBC_TRY(upload_host(xlen.get(), hxl.data(), R * 4, st));
  DevBuf<ScanItem> items;
  uint32_t nitems = 0;
  RBC_TRY(make_bucket_items(qoff.get(), xoff.get(), xlen.get(), R, items, &nitems, ws, 8, st));
  // Scan each group's list through the xMap gather (planKNNKernel / planNNKernel roles).
  TileSrc src{s->dx.mat, s->dx.ld, s->dxMap.mat, nullptr, Q.p, Q.ld, qorder.get(), d, (d + 3) & ~3u};
  if (K == 1) {
    RBC_TRY(fill_f32(od.get(), FLT_MAX, m, st));
    RBC_TRY(fill_u32(oi.get(), kDummy, m, st));
    EpiArgMin epi;
    epi.out_d = od.get();
    epi.out_i = oi.get();
    epi.slot_stride = 0;
    RBC_TRY(launch_tiles_m(metric, src, items.get(), nitems, epi, st));
  } else {
    RBC_TRY(fill_f32(od.get(), FLT_MAX, size_t(m) * K, st));
    RBC_TRY(fill_u32(oi.get(), kDummy, size_t(m) * K, st));
    EpiTopK epi{K, od.get(), oi.get(), K, 0};
    RBC_TRY(launch_tiles_m(metric, src, items.get(), nitems, epi, st));
  }
  RBC_TRY(rows_out(ids, size_t(ldo) * 4, oi.get(), size_t(K) * 4, size_t(K) * 4, m, mem, st));
  if (dists) RBC_TRY(rows_out(dists, size_t(ldo) * 4, od.get(), size_t(K) * 4, size_t(K) * 4, m, mem, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

}  // namespace rbc

using namespace rbc;

extern "C" {

int rbc_build(const float *x, uint32_t n, uint32_t d, uint32_t ldx, uint32_t num_reps, uint32_t s,
              const uint32_t *rep_ids, uint64_t seed, int mem, rbcStruct *out) {
  RBC_ARG_CHECK(out, "NULL rbcStruct");
  RBC_ARG_CHECK(d >= 1 && d <= RBC_MAX_DIM, "dimension %u outside [1, %d]", d, RBC_MAX_DIM);
  RBC_ARG_CHECK(n >= 1, "empty database");
  RBC_ARG_CHECK(num_reps >= 1 && num_reps <= n, "num_reps=%u outside [1, n=%u]", num_reps, n);
  RBC_ARG_CHECK(num_reps <= 32768, "num_reps=%u above the supported 32768", num_reps);
  RBC_ARG_CHECK(s >= 1, "s must be positive");
  hipStream_t st = nullptr;
  const int metric = metric_setting();
  std::vector<uint32_t> reps;
  RBC_TRY(resolve_reps(rep_ids, num_reps, n, seed, reps));
  const uint32_t R = num_reps, pd = pad16(d), pn = pad16(n), pR = pad16(R), ldmap = pad16(s);
  DevBuf<float> dx, dr;
  DevBuf<uint32_t> dmap, drep;
  RBC_TRY(dx.alloc(size_t(pn) * pd));
  RBC_TRY(fill_f32(dx.get(), 0.f, size_t(pn) * pd, st));
  RBC_TRY(rows_in(dx.get(), size_t(pd) * 4, x, size_t(ldx) * 4, size_t(d) * 4, n, mem, st));
  RBC_TRY(dr.alloc(size_t(pR) * pd));
  RBC_TRY(fill_f32(dr.get(), 0.f, size_t(pR) * pd, st));
  RBC_TRY(drep.alloc(R));
  RBC_TRY(upload_host(drep.get(), reps.data(), R * 4, st));
  for (uint32_t r = 0; r < R; ++r)
    RBC_HIP_TRY(hipMemcpyAsync(dr.get() + size_t(r) * pd, dx.get() + size_t(reps[r]) * pd, size_t(pd) * 4,
                               hipMemcpyDeviceToDevice, st));
  RBC_TRY(dmap.alloc(size_t(pR) * ldmap));
  RBC_TRY(fill_u32(dmap.get(), 0, size_t(pR) * ldmap, st));
  unint *gc = static_cast<unint *>(calloc(pR, sizeof(unint)));
  RBC_ARG_CHECK(gc, "host allocation failed");
  struct HostFree {
    unint *p;
    ~HostFree() { free(p); }
  } gcg{gc};

  // Rep chunks: a multiple of 4 rows (the bisection's block) with D <= 2 GiB.
  uint32_t chunk = (uint32_t)std::min<size_t>(R, std::max<size_t>(4, (size_t(2) << 30) / (size_t(n) * 4)));
  chunk = std::max(4u, chunk & ~3u);
  const uint32_t nseg = cdiv(n, kSeg);
  DevBuf<float> D, ranges;
  DevBuf<uint32_t> segc, nonconv, tot;
  RBC_TRY(D.alloc(size_t(chunk) * n));
  RBC_TRY(ranges.alloc(chunk));
  RBC_TRY(segc.alloc(size_t(chunk) * nseg + 1));
  RBC_TRY(nonconv.alloc(1));
  RBC_TRY(tot.alloc(chunk));
  RBC_TRY(fill_u32(nonconv.get(), 0, 1, st));
  Workspace ws;
  for (uint32_t row = 0; row < R; row += chunk) {
    const uint32_t pi = std::min(chunk, R - row);
    // Distance block D[pi][n] (dist1Kernel role): reps as the query side.
    std::vector<ScanItem> iv;
    const uint32_t seg = 1u << 16;
    for (uint32_t x0 = 0; x0 < n; x0 += seg)
      for (uint32_t q0 = 0; q0 < pi; q0 += kTQ)
        iv.push_back(ScanItem{x0, std::min(seg, n - x0), q0, std::min((uint32_t)kTQ, pi - q0), x0});
    DevBuf<ScanItem> items;
    RBC_TRY(upload_items(iv, items, st));
    TileSrc src{dx.get(), pd, nullptr, nullptr, dr.get() + size_t(row) * pd, pd, nullptr, d, (d + 3) & ~3u};
    EpiDense epi{D.get(), n};
    RBC_TRY(launch_tiles_m(metric, src, items.get(), (uint32_t)iv.size(), epi, st));
    hipLaunchKernelGGL(find_range, dim3(cdiv(pi, 4)), dim3(256), 0, st, D.get(), pi, n, s, kMaxBisect,
                       ranges.get(), nonconv.get());
    RBC_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(flag_count, dim3(nseg, pi), dim3(256), 0, st, D.get(), n, ranges.get(), nseg,
                       segc.get());
    RBC_HIP_TRY(hipGetLastError());
    RBC_TRY(exclusive_scan_u32(segc.get(), segc.get(), size_t(pi) * nseg, ws, 0, st));
    hipLaunchKernelGGL(flag_scatter, dim3(nseg, pi), dim3(256), 0, st, D.get(), n, ranges.get(), nseg,
                       segc.get(), row, dmap.get(), ldmap);
    RBC_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(row_totals, dim3(cdiv(pi, 256)), dim3(256), 0, st, segc.get(), pi, nseg, tot.get());
    RBC_HIP_TRY(hipGetLastError());
    RBC_HIP_TRY(hipMemcpyAsync(gc + row, tot.get(), size_t(pi) * 4, hipMemcpyDeviceToHost, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
  }
  uint32_t bad = 0;
  RBC_HIP_TRY(hipMemcpy(&bad, nonconv.get(), 4, hipMemcpyDeviceToHost));
  if (bad) {
    set_error("range bisection did not settle within %u steps for %u rep block(s)", kMaxBisect, bad);
    return RBC_ERR_NOCONV;
  }
  for (uint32_t r = 0; r < R; ++r) gc[r] = std::min(gc[r], ldmap);
  out->dx = matrix{dx.get(), n, d, pn, pd, pd};
  out->dr = matrix{dr.get(), R, d, pR, pd, pd};
  out->dxMap = intMatrix{dmap.get(), R, s, pR, ldmap, ldmap};
  out->groupCount = gc;
  dx.p = dr.p = nullptr;
  dmap.p = nullptr;
  gcg.p = nullptr;
  return RBC_OK;
}

int rbc_query(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs, uint32_t *nn_ids,
              float *nn_dists, int mem, void *stream) {
  if (m == 0) return RBC_OK;
  return oneshot_query_common(q, m, ldq, rbcs, 1, nn_ids, nn_dists, 1, mem, static_cast<hipStream_t>(stream));
}

int rbc_kquery(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs, uint32_t *nn_ids,
               float *nn_dists, uint32_t ldo, int mem, void *stream) {
  RBC_ARG_CHECK(ldo >= KMAX, "output leading dimension %u < KMAX", ldo);
  if (m == 0) return RBC_OK;
  return oneshot_query_common(q, m, ldq, rbcs, KMAX, nn_ids, nn_dists, ldo, mem,
                              static_cast<hipStream_t>(stream));
}

int rbc_query_plan(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs, const uint8_t *plan,
                   uint32_t ldp, uint32_t *nn_ids, float *nn_dists, int mem, void *stream) {
  RBC_ARG_CHECK(rbcs && rbcs->dr.mat, "RBC not built");
  RBC_ARG_CHECK(plan, "NULL plan");
  RBC_ARG_CHECK(ldp >= rbcs->dr.r, "plan leading dimension %u < number of reps %u", ldp, rbcs->dr.r);
  if (m == 0) return RBC_OK;
  return oneshot_query_common(q, m, ldq, rbcs, 1, nn_ids, nn_dists, 1, mem, static_cast<hipStream_t>(stream), plan,
                              ldp);
}

int rbc_kquery_plan(const float *q, uint32_t m, uint32_t ldq, const rbcStruct *rbcs, const uint8_t *plan,
                    uint32_t ldp, uint32_t *nn_ids, float *nn_dists, uint32_t ldo, int mem, void *stream) {
  RBC_ARG_CHECK(rbcs && rbcs->dr.mat, "RBC not built");
  RBC_ARG_CHECK(plan, "NULL plan");
  RBC_ARG_CHECK(ldp >= rbcs->dr.r, "plan leading dimension %u < number of reps %u", ldp, rbcs->dr.r);
  RBC_ARG_CHECK(ldo >= KMAX, "output leading dimension %u < KMAX", ldo);
  if (m == 0) return RBC_OK;
  return oneshot_query_common(q, m, ldq, rbcs, KMAX, nn_ids, nn_dists, ldo, mem, static_cast<hipStream_t>(stream),
                              plan, ldp);
}

int rbc_destroy(rbcStruct *s) {
  RBC_ARG_CHECK(s, "NULL rbcStruct");
  if (s->dx.mat) (void)hipFree(s->dx.mat);
  if (s->dxMap.mat) (void)hipFree(s->dxMap.mat);
  if (s->dr.mat) (void)hipFree(s->dr.mat);
  free(s->groupCount);
  s->dx.mat = nullptr;
  s->dxMap.mat = nullptr;
  s->dr.mat = nullptr;
  s->groupCount = nullptr;
  return RBC_OK;
}

}  // extern "C"
