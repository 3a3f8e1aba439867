// Work split of the persistent list-scan kernels: the list-major (list, query tile)
// items are cut into one contiguous range per XCD, of about equal work, so that the items
// of a list run at the same time on one XCD and re-read the list from its L2.
#pragma once

#include "rbc_device.hpp"

namespace rbc {

constexpr int kXCD = 8;  // XCDs of an MI355X; workgroup b is dispatched to XCD b % kXCD

// Splits the list-major scan items into kXCD contiguous ranges of about equal work and
// clears the per-range counters.  The work of an item is its number of steps times (its
// live waves + 1): the live waves issue the MFMAs, and every item pays the ring loads and
// barriers of all its steps.  A list's items all have the same steps, so a range boundary
// inside a list falls between its full items (VQ / QW waves each).  One block of 1024 threads; items
// are numbered as make_bucket_items numbers them (bucket-major, ceil(count / VQ) each).
template <uint32_t VQ, uint32_t VP, uint32_t QW>
__global__ __launch_bounds__(1024) void scan_split_kernel(const uint32_t *qoff, const uint32_t *len, uint32_t R,
                                                          uint32_t *bounds, uint32_t *next) {
  __shared__ unsigned long long sw[16];
  __shared__ uint32_t si[16];
  __shared__ unsigned long long carry_w;
  __shared__ uint32_t carry_i;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  auto bucket = [&](uint32_t b, unsigned long long &wt, uint32_t &ni, uint32_t &steps) {
    wt = 0;
    ni = 0;
    steps = 0;
    if (b < R) {
      const uint32_t cnt = qoff[b + 1] - qoff[b];
      ni = (cnt + VQ - 1) / VQ;
      steps = (len[b] + VP - 1) / VP;
      wt = (unsigned long long)steps * ((cnt + QW - 1) / QW + ni);
    }
  };
  // Total work.
  unsigned long long tw = 0;
  uint32_t ti = 0;
  for (uint32_t b = t; b < R; b += 1024) {
    unsigned long long wt;
    uint32_t ni, steps;
    bucket(b, wt, ni, steps);
    tw += wt;
    ti += ni;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    tw += __shfl_xor(tw, o, kWave);
    ti += __shfl_xor(ti, o, kWave);
  }
  if (l == 0) {
    sw[w] = tw;
    si[w] = ti;
  }
  __syncthreads();
  unsigned long long W = 0;
  uint32_t NI = 0;
  for (int i = 0; i < 16; ++i) {
    W += sw[i];
    NI += si[i];
  }
  // Boundaries not placed below (no work at all) leave everything to XCD 0's range.
  if (t <= kXCD) bounds[t] = t == 0 ? 0 : NI;
  if (t < kXCD) next[t] = 0;
  if (t == 0) {
    carry_w = 0;
    carry_i = 0;
  }
  __syncthreads();
  // Exclusive prefix of work and items, 1024 buckets at a time; the bucket holding target
  // x W / kXCD places boundary x.
  for (uint32_t c0 = 0; c0 < R; c0 += 1024) {
    unsigned long long wt;
    uint32_t ni, steps;
    bucket(c0 + t, wt, ni, steps);
    unsigned long long pw = wt;
    uint32_t pi = ni;
    for (int o = 1; o < kWave; o <<= 1) {
      const unsigned long long nw = __shfl_up(pw, o, kWave);
      const uint32_t nn = __shfl_up(pi, o, kWave);
      if (l >= o) {
        pw += nw;
        pi += nn;
      }
    }
    __syncthreads();
    if (l == kWave - 1) {
      sw[w] = pw;
      si[w] = pi;
    }
    __syncthreads();
    unsigned long long bw = carry_w;
    uint32_t bi = carry_i;
    for (int i = 0; i < w; ++i) {
      bw += sw[i];
      bi += si[i];
    }
    pw = bw + pw - wt;  // exclusive
    pi = bi + pi - ni;
    if (wt > 0) {
      for (int x = 1; x < kXCD; ++x) {
        const unsigned long long T = W * (unsigned long long)x / kXCD;
        if (T >= pw && T < pw + wt) {
          const unsigned long long full = (unsigned long long)steps * (VQ / QW + 1);
          const uint32_t j = (uint32_t)min((unsigned long long)(ni - 1), (T - pw) / full);
          bounds[x] = pi + j;
        }
      }
    }
    __syncthreads();
    if (t == 1023) {
      carry_w = pw + wt;
      carry_i = pi + ni;
    }
  }
}

}  // namespace rbc
