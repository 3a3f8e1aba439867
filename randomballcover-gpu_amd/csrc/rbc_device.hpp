// Device building blocks shared by the RBC kernels (gfx950, wave64):
//  * (distance, index) pairs with the lowest-index-first total order;
//  * wave-level bitonic sort and merge of 64 such pairs on cross-lane shuffles;
//  * FP32 distance terms in the two summation orders the engine uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace rbc {

constexpr uint32_t kDummy = 0xFFFFFFFFu;
constexpr int kWave = 64;

// Device bounds checks (a build with -DRBC_CHECK).  RBC_CHECKED(i, n) is i when i < n;
// otherwise it records its source line in this translation unit's check word and yields
// 0, an index that is always in bounds, so the kernel finishes normally; the entry point
// then returns RBC_ERR_CHECK naming the translation unit and line (device_check).
// Without RBC_CHECK it is just i.
#ifdef RBC_CHECK
void register_check_reader(int (*read)(unsigned int *line), const char *unit);
namespace {
__device__ unsigned int tu_check_word;
__device__ __noinline__ void check_fail(unsigned int line) { atomicCAS(&tu_check_word, 0u, line); }
// Reads and clears this translation unit's check word (current device).
int read_tu_check(unsigned int *line) {
  unsigned int v = 0, z = 0;
  if (hipMemcpyFromSymbol(&v, HIP_SYMBOL(tu_check_word), sizeof(v)) != hipSuccess) return -1;
  if (v && hipMemcpyToSymbol(HIP_SYMBOL(tu_check_word), &z, sizeof(z)) != hipSuccess) return -1;
  *line = v;
  return 0;
}
struct TuCheckRegistration {
  TuCheckRegistration() { register_check_reader(read_tu_check, __BASE_FILE__); }
} tu_check_registration;
}  // namespace
#define RBC_CHECKED(i, n) ((i) < (n) ? (i) : (::rbc::check_fail(__LINE__), decltype(i)(0)))
#else
#define RBC_CHECKED(i, n) (i)
#endif

// (a, ai) < (b, bi) in the order (distance, index).
__device__ __forceinline__ bool pair_less(float a, uint32_t ai, float b, uint32_t bi) {
  return a < b || (a == b && ai < bi);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// One compare-exchange stage of a bitonic network across lanes at distance j.
// keep_min selects whether this lane keeps the smaller of (own, partner).
__device__ __forceinline__ void cx_stage(float &v, uint32_t &id, int j, bool keep_min) {
  float pv = __shfl_xor(v, j, kWave);
  uint32_t pi = __shfl_xor(id, j, kWave);
  bool take = keep_min ? pair_less(pv, pi, v, id) : pair_less(v, id, pv, pi);
  if (take) {
    v = pv;
    id = pi;
  }
}

// Sorts the 64 pairs held one per lane into ascending (distance, index) order.
__device__ __forceinline__ void wave_bitonic_sort(float &v, uint32_t &id) {
  const int l = lane_id();
#pragma unroll
  for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      bool up = (l & k) == 0;
      bool lower = (l & j) == 0;
      cx_stage(v, id, j, lower == up);
    }
  }
}

// Sorts 64 pairs whose aligned 16-lane runs are already ordered, ascending where lane bit 4
// is clear and descending where it is set (the state wave_bitonic_sort reaches after its
// 16-wide stages): the remaining 32- and 64-wide stages of that network.
__device__ __forceinline__ void wave_bitonic_sort_from16(float &v, uint32_t &id) {
  const int l = lane_id();
#pragma unroll
  for (int k = 32; k <= kWave; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      bool up = (l & k) == 0;
      bool lower = (l & j) == 0;
      cx_stage(v, id, j, lower == up);
    }
  }
}

// Sorts a bitonic sequence held one per lane into ascending order.
__device__ __forceinline__ void wave_bitonic_merge(float &v, uint32_t &id) {
  const int l = lane_id();
#pragma unroll
  for (int j = kWave >> 1; j > 0; j >>= 1) cx_stage(v, id, j, (l & j) == 0);
}

// a (ascending, one per lane) and b (ascending, one per lane): returns in (a) the 64
// smallest pairs of the union, ascending.
__device__ __forceinline__ void wave_merge_lower(float &av, uint32_t &ai, float bv, uint32_t bi) {
  const int l = lane_id();
  float rv = __shfl(bv, kWave - 1 - l, kWave);
  uint32_t ri = __shfl(bi, kWave - 1 - l, kWave);
  if (pair_less(rv, ri, av, ai)) {
    av = rv;
    ai = ri;
  }
  wave_bitonic_merge(av, ai);
}

// Per-coordinate distance term.
template <int METRIC>
__device__ __forceinline__ float fma_term(float x, float q, float acc) {
  float t = x - q;
  if constexpr (METRIC == 0) {
    return __builtin_fmaf(t, t, acc);
  } else {
    return acc + __builtin_fabsf(t);
  }
}

// Tree-order term (product rounded before any addition).
template <int METRIC>
__device__ __forceinline__ float tree_term(float x, float q) {
  float t = x - q;
  if constexpr (METRIC == 0) {
    return __fmul_rn(t, t);
  } else {
    return __builtin_fabsf(t);
  }
}

// Butterfly sum across the wave: t[i] = t[i] + t[i ^ s] for s = 32..1.  Lane 0's
// result equals the stride-halving tree over lanes 0..63.
__device__ __forceinline__ float wave_tree_sum(float v) {
#pragma unroll
  for (int s = kWave / 2; s > 0; s >>= 1) v = __fadd_rn(v, __shfl_xor(v, s, kWave));
  return v;
}

}  // namespace rbc
