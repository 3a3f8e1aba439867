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

// One planned list of a query group: its xMap entries [x_begin, x_begin + x_count).
struct PlanList {
  uint32_t x_begin, x_count;
};

// Queries src.qids[hqoff[g] .. hqoff[g+1]) (group g) scan lists[g] in order; m rows of K
// results (ascending (distance, id)) land in od/oi on the device.
static int oneshot_query_plan(const TileSrc &src, uint32_t m, uint32_t K, const std::vector<uint32_t> &hqoff,
                              const std::vector<std::vector<PlanList>> &lists, int metric, float *od, uint32_t *oi,
                              hipStream_t st) {
  size_t most = 0;
  for (const auto &l : lists) most = std::max(most, l.size());
  const uint32_t per = kPlanSlots - 1;
  const size_t slot = size_t(m) * K;
  DevBuf<float> pd;
  DevBuf<uint32_t> pi;
  RBC_TRY(pd.alloc(kPlanSlots * slot));
  RBC_TRY(pi.alloc(kPlanSlots * slot));
  RBC_TRY(fill_f32(pd.get(), FLT_MAX, slot, st));
  RBC_TRY(fill_u32(pi.get(), kDummy, slot, st));
  DevBuf<ScanItem> items;
  for (size_t first = 0; first < most; first += per) {
    std::vector<ScanItem> iv;
    for (size_t g = 0; g < lists.size(); ++g) {
      const uint32_t cnt = hqoff[g + 1] - hqoff[g];
      for (size_t k = first; k < std::min(first + per, lists[g].size()); ++k) {
        if (lists[g][k].x_count == 0) continue;
        for (uint32_t t = 0; t < cnt; t += kTQ) {
          ScanItem it;
          it.x_begin = lists[g][k].x_begin;
          it.x_count = lists[g][k].x_count;
          it.q_begin = hqoff[g] + t;
          it.q_count = std::min<uint32_t>(kTQ, cnt - t);
          it.aux = 1 + (uint32_t)(k - first);
          iv.push_back(it);
        }
      }
    }
    RBC_TRY(fill_f32(pd.get() + slot, FLT_MAX, per * slot, st));
    RBC_TRY(fill_u32(pi.get() + slot, kDummy, per * slot, st));
    if (!iv.empty()) {
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
    // Planned lists of each group that has queries, ascending j.
    std::vector<std::vector<PlanList>> lists(R);
    for (uint32_t i = 0; i < R; ++i) {
      if (hqoff[i + 1] == hqoff[i]) continue;
      for (uint32_t j = 0; j < R; ++j)
        if (plan[size_t(i) * ldp + j]) lists[i].push_back(PlanList{hxo[j], hxl[j]});
    }
    TileSrc src{s->dx.mat, s->dx.ld, s->dxMap.mat, nullptr, Q.p, Q.ld, qorder.get(), d, (d + 3) & ~3u};
    RBC_TRY(oneshot_query_plan(src, m, K, hqoff, lists, metric, od.get(), oi.get(), st));
    RBC_TRY(rows_out(ids, size_t(ldo) * 4, oi.get(), size_t(K) * 4, size_t(K) * 4, m, mem, st));
    if (dists) RBC_TRY(rows_out(dists, size_t(ldo) * 4, od.get(), size_t(K) * 4, size_t(K) * 4, m, mem, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
    return device_check();
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
  return device_check();
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

// ------------------------------------------------ the reference's internals (rbc.h:15-26)

int rbc_dist_sub(const float *r, uint32_t ldr, uint32_t start, uint32_t length, const float *x, uint32_t nx,
                 uint32_t ldx, uint32_t d, float *D, uint32_t ldd, void *stream) {
  RBC_ARG_CHECK(d >= 1 && d <= RBC_MAX_DIM, "dimension %u outside [1, %d]", d, RBC_MAX_DIM);
  RBC_ARG_CHECK(ldr >= d && ldx >= d && ldr % 4 == 0 && ldx % 4 == 0,
                "leading dimensions (%u, %u) must be >= d = %u and multiples of 4", ldr, ldx, d);
  RBC_ARG_CHECK(ldd >= nx, "distance matrix leading dimension %u < %u points", ldd, nx);
  if (!length || !nx) return RBC_OK;
  RBC_ARG_CHECK(r && x && D, "NULL pointer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<ScanItem> iv;
  const uint32_t seg = 1u << 16;
  for (uint32_t x0 = 0; x0 < nx; x0 += seg)
    for (uint32_t q0 = 0; q0 < length; q0 += kTQ)
      iv.push_back(ScanItem{x0, std::min(seg, nx - x0), q0, std::min((uint32_t)kTQ, length - q0), x0});
  DevBuf<ScanItem> items;
  RBC_TRY(upload_items(iv, items, st));
  TileSrc src{x, ldx, nullptr, nullptr, r + size_t(start) * ldr, ldr, nullptr, d, (d + 3) & ~3u};
  EpiDense epi{D, ldd};
  RBC_TRY(launch_tiles_m(metric_setting(), src, items.get(), (uint32_t)iv.size(), epi, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int rbc_compute_reps(const float *q, uint32_t m, uint32_t ldq, const float *reps, uint32_t R, uint32_t ldr,
                     uint32_t d, uint32_t *rep_ids, float *dist_to_reps, int mem_in, int mem_out, void *stream) {
  RBC_ARG_CHECK(d >= 1 && d <= RBC_MAX_DIM, "dimension %u outside [1, %d]", d, RBC_MAX_DIM);
  RBC_ARG_CHECK(R >= 1, "no representatives");
  RBC_ARG_CHECK(rep_ids && dist_to_reps, "NULL output");
  if (m == 0) return RBC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevMat Q, Rm;
  RBC_TRY(Q.stage(q, m, d, ldq, mem_in, st));
  RBC_TRY(Rm.stage(reps, R, d, ldr, mem_in, st));
  DevBuf<uint32_t> ni;
  DevBuf<float> nd;
  RBC_TRY(ni.alloc(m));
  RBC_TRY(nd.alloc(m));
  RBC_TRY(nn_device(Rm.p, Rm.ld, R, nullptr, Q.p, Q.ld, m, d, metric_setting(), nd.get(), ni.get(), st));
  RBC_TRY(rows_out(rep_ids, 4, ni.get(), 4, 4, m, mem_out, st));
  RBC_TRY(rows_out(dist_to_reps, 4, nd.get(), 4, 4, m, mem_out, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int rbc_plan_init(compPlan *out, const char *cm, uint32_t ldcm, const uint32_t *group_count_q,
                  const uint32_t *group_count_x, uint32_t num_reps) {
  RBC_ARG_CHECK(out && group_count_q && group_count_x && (cm || num_reps == 0), "NULL argument");
  RBC_ARG_CHECK(ldcm >= num_reps, "computation matrix leading dimension %u < %u reps", ldcm, num_reps);
  const uint32_t R = num_reps;
  std::vector<uint32_t> ng(R, 0);
  uint32_t ld = 0;
  for (uint32_t i = 0; i < R; ++i) {
    for (uint32_t j = 0; j < R; ++j) ng[i] += cm[size_t(i) * ldcm + j] != 0;
    ld = std::max(ld, ng[i]);
  }
  std::vector<uint32_t> q2g;
  for (uint32_t i = 0; i < R; ++i) q2g.insert(q2g.end(), pad16(group_count_q[i]), i);
  std::vector<uint32_t> g2x(size_t(R) * ld, 0), gcx(size_t(R) * ld, 0);
  for (uint32_t i = 0; i < R; ++i)
    for (uint32_t j = 0, k = 0; j < R; ++j)
      if (cm[size_t(i) * ldcm + j]) {
        g2x[size_t(i) * ld + k] = j;
        gcx[size_t(i) * ld + k++] = group_count_x[j];
      }
  DevBuf<uint32_t> dng, dgcx, dq2g, dg2x;
  RBC_TRY(dng.alloc(std::max<size_t>(R, 1)));
  RBC_TRY(dgcx.alloc(std::max<size_t>(gcx.size(), 1)));
  RBC_TRY(dq2g.alloc(std::max<size_t>(q2g.size(), 1)));
  RBC_TRY(dg2x.alloc(std::max<size_t>(g2x.size(), 1)));
  RBC_TRY(upload_host(dng.get(), ng.data(), ng.size() * 4, nullptr));
  RBC_TRY(upload_host(dgcx.get(), gcx.data(), gcx.size() * 4, nullptr));
  RBC_TRY(upload_host(dq2g.get(), q2g.data(), q2g.size() * 4, nullptr));
  RBC_TRY(upload_host(dg2x.get(), g2x.data(), g2x.size() * 4, nullptr));
  out->numGroups = dng.p;
  out->groupCountX = dgcx.p;
  out->qToQGroup = dq2g.p;
  out->qGroupToXGroup = dg2x.p;
  out->ld = ld;
  dng.p = dgcx.p = dq2g.p = dg2x.p = nullptr;
  return RBC_OK;
}

int rbc_plan_free(compPlan *p) {
  RBC_ARG_CHECK(p, "NULL compPlan");
  for (unint **a : {&p->numGroups, &p->groupCountX, &p->qToQGroup, &p->qGroupToXGroup}) {
    if (*a) (void)hipFree(*a);
    *a = nullptr;
  }
  return RBC_OK;
}

int rbc_plan_query(const float *x, uint32_t nx, uint32_t ldx, const uint32_t *xmap, uint32_t xmap_rows,
                   uint32_t ldxmap, const float *q, uint32_t m, uint32_t ldq, uint32_t d, const uint32_t *qmap,
                   uint32_t comp_length, const compPlan *plan, uint32_t K, uint32_t *ids, float *dists, uint32_t ldo,
                   int mem_out, void *stream) {
  RBC_ARG_CHECK(plan, "NULL compPlan");
  RBC_ARG_CHECK(d >= 1 && d <= RBC_MAX_DIM, "dimension %u outside [1, %d]", d, RBC_MAX_DIM);
  RBC_ARG_CHECK(K >= 1 && K <= KMAX, "K=%u outside [1, %d]", K, KMAX);
  RBC_ARG_CHECK(ldo >= K, "output leading dimension %u < K=%u", ldo, K);
  RBC_ARG_CHECK(ids && (dists || K == 1), "NULL output");
  RBC_ARG_CHECK(ldx >= d && ldx % 4 == 0, "database leading dimension %u must be >= d and a multiple of 4", ldx);
  if (m == 0) return RBC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // The plan's query slots run in blocks of 16 that all take the group of the block's
  // first slot (planNNKernel, kernels.cu:32-33); a partial last block is not run
  // (planNNWrap, kernelWrap.cu:121-126).
  const uint32_t slots = comp_length / BLOCK_SIZE * BLOCK_SIZE;
  std::vector<uint32_t> hq(slots), hg(slots);
  if (slots) {
    RBC_ARG_CHECK(qmap && plan->qToQGroup, "NULL query map");
    RBC_HIP_TRY(hipMemcpy(hq.data(), qmap, size_t(slots) * 4, hipMemcpyDeviceToHost));
    RBC_HIP_TRY(hipMemcpy(hg.data(), plan->qToQGroup, size_t(slots) * 4, hipMemcpyDeviceToHost));
  }
  uint32_t G = 0;
  for (uint32_t b = 0; b < slots; b += BLOCK_SIZE) G = std::max(G, hg[b] + 1);
  const uint32_t ld = plan->ld;
  std::vector<uint32_t> hng(G), hg2x(size_t(G) * ld), hgcx(size_t(G) * ld);
  if (G) {
    RBC_HIP_TRY(hipMemcpy(hng.data(), plan->numGroups, size_t(G) * 4, hipMemcpyDeviceToHost));
    if (ld) {
      RBC_HIP_TRY(hipMemcpy(hg2x.data(), plan->qGroupToXGroup, hg2x.size() * 4, hipMemcpyDeviceToHost));
      RBC_HIP_TRY(hipMemcpy(hgcx.data(), plan->groupCountX, hgcx.size() * 4, hipMemcpyDeviceToHost));
    }
  }
  // Queries of each group (slot order), each query at most once.
  std::vector<std::vector<uint32_t>> members(G);
  std::vector<char> seen(m, 0);
  for (uint32_t b = 0; b < slots; b += BLOCK_SIZE)
    for (uint32_t s = b; s < b + BLOCK_SIZE; ++s) {
      const uint32_t qi = hq[s];
      if (qi == DUMMY_IDX) continue;
      RBC_ARG_CHECK(qi < m, "query map entry %u at slot %u >= %u queries", qi, s, m);
      RBC_ARG_CHECK(!seen[qi], "query %u appears twice in the query map", qi);
      seen[qi] = 1;
      members[hg[b]].push_back(qi);
    }
  // Planned lists of each group, checked against the xMap's extent and contents.
  std::vector<uint32_t> hxmap;
  if (G && ld) {
    RBC_ARG_CHECK(xmap, "NULL xMap");
    hxmap.resize(size_t(xmap_rows) * ldxmap);
    if (!hxmap.empty())
      RBC_HIP_TRY(hipMemcpy(hxmap.data(), xmap, hxmap.size() * 4, hipMemcpyDeviceToHost));
  }
  std::vector<std::vector<PlanList>> lists(G);
  for (uint32_t g = 0; g < G; ++g) {
    if (members[g].empty()) continue;
    RBC_ARG_CHECK(hng[g] <= ld, "group %u plans %u lists, more than the plan's ld %u", g, hng[g], ld);
    for (uint32_t k = 0; k < hng[g]; ++k) {
      const uint32_t xg = hg2x[size_t(g) * ld + k], cnt = hgcx[size_t(g) * ld + k];
      RBC_ARG_CHECK(xg < xmap_rows, "group %u plans list %u of %u", g, xg, xmap_rows);
      RBC_ARG_CHECK(cnt <= ldxmap, "list %u count %u exceeds the xMap row length %u", xg, cnt, ldxmap);
      for (uint32_t p = 0; p < cnt; ++p)
        RBC_ARG_CHECK(hxmap[size_t(xg) * ldxmap + p] < nx, "xMap[%u][%u] = %u >= %u points", xg, p,
                      hxmap[size_t(xg) * ldxmap + p], nx);
      lists[g].push_back(PlanList{xg * ldxmap, cnt});
    }
  }
  std::vector<uint32_t> order, hqoff(G + 1, 0);
  for (uint32_t g = 0; g < G; ++g) {
    order.insert(order.end(), members[g].begin(), members[g].end());
    hqoff[g + 1] = (uint32_t)order.size();
  }
  DevMat Q;
  RBC_TRY(Q.stage(q, m, d, ldq, RBC_MEM_DEVICE, st));
  DevBuf<uint32_t> qorder, oi;
  DevBuf<float> od;
  RBC_TRY(qorder.alloc(std::max<size_t>(order.size(), 1)));
  if (!order.empty()) RBC_TRY(upload_host(qorder.get(), order.data(), order.size() * 4, st));
  RBC_TRY(od.alloc(size_t(m) * K));
  RBC_TRY(oi.alloc(size_t(m) * K));
  TileSrc src{x, ldx, xmap, nullptr, Q.p, Q.ld, qorder.get(), d, (d + 3) & ~3u};
  RBC_TRY(oneshot_query_plan(src, m, K, hqoff, lists, metric_setting(), od.get(), oi.get(), st));
  RBC_TRY(rows_out(ids, size_t(ldo) * 4, oi.get(), size_t(K) * 4, size_t(K) * 4, m, mem_out, st));
  if (dists) RBC_TRY(rows_out(dists, size_t(ldo) * 4, od.get(), size_t(K) * 4, size_t(K) * 4, m, mem_out, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int rbc_setup_reps(const float *x, uint32_t n, uint32_t d, uint32_t ldx, const uint32_t *rep_ids, uint32_t R,
                   matrix *out) {
  RBC_ARG_CHECK(out && x && rep_ids, "NULL argument");
  RBC_ARG_CHECK(ldx >= d && d >= 1, "leading dimension %u < dimension %u", ldx, d);
  RBC_ARG_CHECK(R >= 1 && R <= n, "num_reps=%u outside [1, n=%u]", R, n);
  const uint32_t pd = pad16(d), pR = pad16(R);
  std::vector<float> h(size_t(pR) * pd, 0.f);
  for (uint32_t i = 0; i < R; ++i) {
    RBC_ARG_CHECK(rep_ids[i] < n, "rep id %u >= n=%u", rep_ids[i], n);
    std::copy(x + size_t(rep_ids[i]) * ldx, x + size_t(rep_ids[i]) * ldx + d, h.begin() + size_t(i) * pd);
  }
  DevBuf<float> dr;
  RBC_TRY(dr.alloc(h.size()));
  RBC_TRY(upload_host(dr.get(), h.data(), h.size() * 4, nullptr));
  *out = matrix{dr.p, R, d, pR, pd, pd};
  dr.p = nullptr;
  return RBC_OK;
}

}  // extern "C"

