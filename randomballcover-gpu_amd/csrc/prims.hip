// Exclusive scan, stable counting sort and work-item generation (see prims.hpp).
//
// The scan is the work the reference spreads over sumKernel / sumKernelI /
// combineSumKernel (sKernel.cu:14-133, sKernelWrap.cu:46-104): a two-level scheme here,
// with 2048-element blocks scanned in LDS by a wave-level inclusive scan, the block
// totals scanned recursively, and the offsets added back.
#include "prims.hpp"

namespace rbc {

namespace {

constexpr int kScanNT = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanBlock = kScanNT * kScanPerThread;  // 2048 elements per block

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int s = 1; s < kWave; s <<= 1) {
    uint32_t o = __shfl_up(v, s, kWave);
    if (l >= s) v += o;
  }
  return v;
}

// Block-level exclusive scan of kScanBlock elements held 8 per thread (thread-major).
// Returns the block total; vals are replaced by their exclusive prefixes.
__device__ uint32_t block_excl_scan(uint32_t (&vals)[kScanPerThread], uint32_t *lds) {
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i) {
    uint32_t v = vals[i];
    vals[i] = s;
    s += v;
  }
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  uint32_t inc = wave_incl_scan(s);
  if (l == kWave - 1) lds[w] = inc;
  __syncthreads();
  uint32_t wbase = 0, total = 0;
  for (int i = 0; i < kScanNT / kWave; ++i) {
    uint32_t wt = lds[i];
    if (i < w) wbase += wt;
    total += wt;
  }
  uint32_t base = wbase + inc - s;
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i) vals[i] += base;
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kScanNT) void scan_blocks(const uint32_t *in, uint32_t *out,
                                                       size_t n, uint32_t *block_tot) {
  __shared__ uint32_t lds[kScanNT / kWave];
  size_t base = size_t(blockIdx.x) * kScanBlock + size_t(threadIdx.x) * kScanPerThread;
  uint32_t v[kScanPerThread];
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i) v[i] = base + i < n ? in[base + i] : 0u;
  uint32_t tot = block_excl_scan(v, lds);
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i)
    if (base + i < n) out[base + i] = v[i];
  if (threadIdx.x == 0) block_tot[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanNT) void scan_add(uint32_t *out, size_t n, const uint32_t *block_off) {
  size_t base = size_t(blockIdx.x) * kScanBlock + size_t(threadIdx.x) * kScanPerThread;
  uint32_t add = block_off[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i)
    if (base + i < n) out[base + i] += add;
}

__global__ void single_block_off(uint32_t *off, const uint32_t *tot) {
  off[0] = 0;
  off[1] = tot[0];
}

__global__ void write_total(uint32_t *out, size_t n, const uint32_t *block_off, size_t nb) {
  // out[n] = total = block_off[nb] (exclusive scan of block totals has nb+1 entries).
  out[n] = block_off[nb];
}

// ---- counting sort ----
constexpr int kSortNT = 256;
constexpr int kSortItems = 4096;  // items per block

__global__ __launch_bounds__(kSortNT) void sort_hist(const uint32_t *key, size_t n, uint32_t nb_buckets,
                                                     uint32_t nblocks, uint32_t *hist) {
  extern __shared__ uint32_t cnt[];
  for (uint32_t b = threadIdx.x; b < nb_buckets; b += kSortNT) cnt[b] = 0;
  __syncthreads();
  size_t i0 = size_t(blockIdx.x) * kSortItems;
  for (uint32_t j = threadIdx.x; j < kSortItems; j += kSortNT) {
    size_t i = i0 + j;
    if (i < n && key[i] < nb_buckets) atomicAdd(&cnt[key[i]], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nb_buckets; b += kSortNT)
    hist[size_t(b) * nblocks + blockIdx.x] = cnt[b];
}

// Stable scatter of one block's kSortItems keys, 256 at a time.  Each lane finds the lanes
// of its wave holding the same key with one ballot per key bit (kbits bits, plus the
// valid flag); its rank is the number of such lanes below it.  The block's 4 waves then
// take positions from the per-bucket counters in wave order, so equal keys keep their
// input order.
__global__ __launch_bounds__(kSortNT) void sort_scatter(const uint32_t *key, size_t n, uint32_t nb_buckets,
                                                        uint32_t kbits, uint32_t nblocks,
                                                        const uint32_t *hist_off, uint32_t *order) {
  static_assert(kSortNT == 4 * 64, "four waves per block");
  extern __shared__ uint32_t run[];  // [nb_buckets] running positions
  for (uint32_t b = threadIdx.x; b < nb_buckets; b += kSortNT)
    run[b] = hist_off[size_t(b) * nblocks + blockIdx.x];
  __syncthreads();
  const size_t i0 = size_t(blockIdx.x) * kSortItems;
  const uint32_t t = threadIdx.x, w = t / 64, l = t % 64;
  const uint64_t below = l ? (~0ull >> (64 - l)) : 0ull;
  for (uint32_t c = 0; c < kSortItems; c += kSortNT) {
    const size_t i = i0 + c + t;
    const uint32_t k = i < n ? key[i] : 0xFFFFFFFFu;
    const bool valid = k < nb_buckets;
    uint64_t same = __ballot(valid);  // only valid lanes' masks are used
    for (uint32_t b = 0; b < kbits; ++b) {
      const bool bit = (k >> b) & 1u;
      const uint64_t set = __ballot(bit);
      same &= bit ? set : ~set;
    }
    const uint32_t rank = (uint32_t)__popcll(same & below);
    const bool leader = valid && (same & below) == 0;
    for (uint32_t ww = 0; ww < 4; ++ww) {
      if (w == ww && valid) {
        const uint32_t base = run[k];
        order[RBC_CHECKED(base + rank, (uint32_t)n)] = (uint32_t)i;
        // Every lane of the wave has read run[k] before its leader's write (in-order LDS).
        if (leader) run[k] = base + (uint32_t)__popcll(same);
      }
      __syncthreads();
    }
  }
}

__global__ void gather_offsets(const uint32_t *hist_off, uint32_t nb_buckets, uint32_t nblocks,
                               uint32_t *offsets) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb_buckets) offsets[b] = hist_off[size_t(b) * nblocks];
  if (b == nb_buckets) offsets[b] = hist_off[size_t(nb_buckets) * nblocks];  // items with a valid key
}

// ---- items ----
__global__ void bucket_tiles(const uint32_t *qoff, uint32_t nb, uint32_t tile, uint32_t *ntiles) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) ntiles[b] = (qoff[b + 1] - qoff[b] + tile - 1) / tile;
}

__global__ void bucket_items(const uint32_t *qoff, const uint32_t *xoff, const uint32_t *xlen,
                             uint32_t nb, uint32_t tile, const uint32_t *toff, ScanItem *items) {
  uint32_t b = blockIdx.x;
  if (b >= nb) return;
  uint32_t q0 = qoff[b], cnt = qoff[b + 1] - q0, t0 = toff[b];
  uint32_t nt = (cnt + tile - 1) / tile;
  for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x) {
    ScanItem it;
    it.x_begin = xoff[b];
    it.x_count = xlen[b];
    it.q_begin = q0 + t * tile;
    it.q_count = min(tile, cnt - t * tile);
    it.aux = b;
    items[t0 + t] = it;
  }
}

}  // namespace

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, Workspace &ws, size_t slot,
                       hipStream_t st) {
  size_t nb = cdivz(n, kScanBlock);
  if (nb == 0) nb = 1;
  void *tot_v, *off_v;
  RBC_TRY(ws.get(slot, (nb + 1) * sizeof(uint32_t), &tot_v));
  RBC_TRY(ws.get(slot + 1, (nb + 1) * sizeof(uint32_t), &off_v));
  uint32_t *tot = static_cast<uint32_t *>(tot_v), *off = static_cast<uint32_t *>(off_v);
  hipLaunchKernelGGL(scan_blocks, dim3((uint32_t)nb), dim3(kScanNT), 0, st, in, out, n, tot);
  RBC_HIP_TRY(hipGetLastError());
  if (nb == 1) {
    hipLaunchKernelGGL(single_block_off, dim3(1), dim3(1), 0, st, off, tot);
    RBC_HIP_TRY(hipGetLastError());
  } else {
    RBC_TRY(exclusive_scan_u32(tot, off, nb, ws, slot + 2, st));
    hipLaunchKernelGGL(scan_add, dim3((uint32_t)nb), dim3(kScanNT), 0, st, out, n, off);
    RBC_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(write_total, dim3(1), dim3(1), 0, st, out, n, off, nb);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int counting_sort(const uint32_t *key, size_t n, uint32_t nbuckets, uint32_t *order,
                  uint32_t *offsets, Workspace &ws, size_t slot, hipStream_t st) {
  if (nbuckets > 32768) {
    set_error("counting_sort: %u buckets exceed the 32768 supported", nbuckets);
    return RBC_ERR_LIMIT;
  }
  uint32_t nblocks = (uint32_t)cdivz(n, kSortItems);
  if (nblocks == 0) nblocks = 1;
  size_t hn = size_t(nbuckets) * nblocks;
  void *hv;
  RBC_TRY(ws.get(slot, (hn + 1) * sizeof(uint32_t), &hv));
  uint32_t *hist = static_cast<uint32_t *>(hv);
  hipLaunchKernelGGL(sort_hist, dim3(nblocks), dim3(kSortNT), nbuckets * sizeof(uint32_t), st, key,
                     n, nbuckets, nblocks, hist);
  RBC_HIP_TRY(hipGetLastError());
  RBC_TRY(exclusive_scan_u32(hist, hist, hn, ws, slot + 1, st));
  uint32_t kbits = 0;
  while ((1u << kbits) < nbuckets) ++kbits;
  hipLaunchKernelGGL(sort_scatter, dim3(nblocks), dim3(kSortNT), nbuckets * sizeof(uint32_t), st, key, n,
                     nbuckets, kbits, nblocks, hist, order);
  RBC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(gather_offsets, dim3(cdiv(nbuckets + 1, 256)), dim3(256), 0, st, hist, nbuckets,
                     nblocks, offsets);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int make_bucket_items(const uint32_t *qoff, const uint32_t *xoff, const uint32_t *xlen,
                      uint32_t nbuckets, DevBuf<ScanItem> &items_buf, uint32_t *nitems,
                      Workspace &ws, size_t slot, hipStream_t st, uint32_t tile) {
  void *tv;
  RBC_TRY(ws.get(slot, (size_t(nbuckets) + 1) * sizeof(uint32_t), &tv));
  uint32_t *toff = static_cast<uint32_t *>(tv);
  hipLaunchKernelGGL(bucket_tiles, dim3(cdiv(nbuckets, 256)), dim3(256), 0, st, qoff, nbuckets, tile, toff);
  RBC_HIP_TRY(hipGetLastError());
  RBC_TRY(exclusive_scan_u32(toff, toff, nbuckets, ws, slot + 1, st));
  uint32_t total = 0;
  RBC_HIP_TRY(hipMemcpyAsync(&total, toff + nbuckets, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  *nitems = total;
  if (total == 0) return RBC_OK;
  RBC_TRY(items_buf.ensure(total));
  hipLaunchKernelGGL(bucket_items, dim3(nbuckets), dim3(64), 0, st, qoff, xoff, xlen, nbuckets, tile, toff,
                     items_buf.get());
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

std::vector<ScanItem> grid_items(uint32_t m, uint32_t n, uint32_t seg_len) {
  std::vector<ScanItem> v;
  if (seg_len == 0) seg_len = n ? n : 1;
  uint32_t nseg = n ? cdiv(n, seg_len) : 1;
  for (uint32_t s = 0; s < nseg; ++s)
    for (uint32_t q0 = 0; q0 < m; q0 += kTQ) {
      ScanItem it;
      it.x_begin = s * seg_len;
      it.x_count = n ? std::min(seg_len, n - s * seg_len) : 0;
      it.q_begin = q0;
      it.q_count = std::min((uint32_t)kTQ, m - q0);
      it.aux = s;
      v.push_back(it);
    }
  return v;
}

}  // namespace rbc
