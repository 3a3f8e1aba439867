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

// Cross-lane exchange without the LDS crossbar: the value of lane l ^ J for J = 1..32, on
// DPP (quad permutes for 1 and 2, row shifts for 4, a row rotate for 8) and on gfx950's
// v_permlane16_swap / v_permlane32_swap for 16 and 32 (each swap leaves the partner row's
// value in one of its two results).  These replace __shfl_xor (ds_bpermute_b32) in the
// bitonic sort/merge stages of the top-k and in the butterfly reductions.
template <int J>
__device__ __forceinline__ uint32_t xor_lane_u32(uint32_t v) {
  static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "lane distance");
  if constexpr (J == 1) {
    return __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (J == 2) {
    return __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (J == 4) {
    const uint32_t up = __builtin_amdgcn_update_dpp(0u, v, 0x104, 0xF, 0xF, false);  // row_shl:4, lane l+4
    const uint32_t dn = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, false);  // row_shr:4, lane l-4
    return (lane_id() & 4) ? dn : up;
  } else if constexpr (J == 8) {
    return __builtin_amdgcn_update_dpp(0u, v, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (J == 16) {
    const auto s = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane_id() & 16) ? s[0] : s[1];
  } else {
    const auto s = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane_id() & 32) ? s[0] : s[1];
  }
}

// Run-time distance (a constant once the caller's loop is unrolled).
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, int j) {
  switch (j) {
    case 1: return xor_lane_u32<1>(v);
    case 2: return xor_lane_u32<2>(v);
    case 4: return xor_lane_u32<4>(v);
    case 8: return xor_lane_u32<8>(v);
    case 16: return xor_lane_u32<16>(v);
    default: return xor_lane_u32<32>(v);
  }
}
__device__ __forceinline__ float xor_lane(float v, int j) {
  return __uint_as_float(xor_lane(__float_as_uint(v), j));
}
__device__ __forceinline__ double xor_lane(double v, int j) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = xor_lane((uint32_t)u, j), hi = xor_lane((uint32_t)(u >> 32), j);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// One compare-exchange stage of a bitonic network across lanes at distance j.
// keep_min selects whether this lane keeps the smaller of (own, partner).
__device__ __forceinline__ void cx_stage(float &v, uint32_t &id, int j, bool keep_min) {
  float pv = xor_lane(v, j);
  uint32_t pi = xor_lane(id, j);
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
  for (int s = kWave / 2; s > 0; s >>= 1) v = __fadd_rn(v, xor_lane(v, s));
  return v;
}

}  // namespace rbc
