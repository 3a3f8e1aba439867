// Measured compute peaks of the device, for roofline reports: kernels that issue nothing
// but independent accumulation chains in registers, timed with hipEvents.
//  * BF16 MFMA: v_mfma_f32_16x16x32_bf16 (the list scan's instruction), 16*16*32*2 FLOP
//    per wave instruction.
//  * FP32 FMA: v_fma_f32 (the instruction of the FP32 distance kernels; packed FP32 is
//    disabled in this build), 64*2 FLOP per wave instruction.
#include <hip/hip_runtime.h>

#include "rbc_common.hpp"
#include "rbc_hip.h"

namespace rbc {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kMfmaChains = 8;  // independent accumulators per wave
constexpr int kFmaChains = 16;  // independent accumulators per lane
constexpr int kBlock = 256;
constexpr int kBlocksPerCU = 8;  // 8 waves per SIMD

__global__ __launch_bounds__(kBlock) void mfma_bf16_loop(uint32_t iters, float *out) {
  bf16x8 a, b;
  for (int e = 0; e < 8; ++e) {
    a[e] = __bf16(float((threadIdx.x + e) & 7) * 0.125f);
    b[e] = __bf16(float((threadIdx.x * 3 + e) & 7) * 0.0625f);
  }
  // Built with the VGPR form of MFMA: the loop body is the eight MFMAs and the counter.
  f32x4 acc[kMfmaChains];
#pragma unroll
  for (int c = 0; c < kMfmaChains; ++c) acc[c] = f32x4{float(c), 0.f, 0.f, 0.f};
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kMfmaChains; ++c) acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
  }
  f32x4 t = acc[0];
#pragma unroll
  for (int c = 1; c < kMfmaChains; ++c) t += acc[c];
  const float s = t[0] + t[1] + t[2] + t[3];
  out[size_t(blockIdx.x) * kBlock + threadIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void fma_f32_loop(uint32_t iters, float *out) {
  float acc[kFmaChains];
  for (int c = 0; c < kFmaChains; ++c) acc[c] = float(threadIdx.x + c) * 1e-3f;
  const float mul = 0.999f, add = float(blockIdx.x & 15) * 1e-6f;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kFmaChains; ++c) acc[c] = fmaf(acc[c], mul, add);
  }
  float s = 0.f;
  for (int c = 0; c < kFmaChains; ++c) s += acc[c];
  out[size_t(blockIdx.x) * kBlock + threadIdx.x] = s;
}

// Runs `launch(iters)` once untimed and once timed; returns the timed milliseconds.
template <class F>
int time_launch(F launch, uint32_t iters, hipStream_t st, float *ms) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RBC_HIP_TRY(hipEventCreate(&e0));
  hipError_t err = hipEventCreate(&e1);
  if (err == hipSuccess) {
    launch(iters / 16 + 1);
    err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e0, st);
    if (err == hipSuccess) {
      launch(iters);
      err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipEventRecord(e1, st);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
  }
  (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (err != hipSuccess) {
    set_error("peak measurement failed: %s", hipGetErrorString(err));
    return RBC_ERR_HIP;
  }
  return RBC_OK;
}

}  // namespace
}  // namespace rbc

extern "C" int rbc_measure_peaks(double *bf16_mfma_tflops, double *fp32_fma_tflops, void *stream) {
  using namespace rbc;
  RBC_ARG_CHECK(bf16_mfma_tflops && fp32_fma_tflops, "rbc_measure_peaks: null output");
  hipStream_t st = static_cast<hipStream_t>(stream);
  int dev = 0, cus = 0;
  RBC_HIP_TRY(hipGetDevice(&dev));
  RBC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t grid = uint32_t(cus) * kBlocksPerCU;
  DevBuf<float> out;
  RBC_TRY(out.alloc(size_t(grid) * kBlock));
  const double waves = double(grid) * (kBlock / 64);

  // About 20 ms each at the published MI355X rates.
  const uint32_t it_mfma = 40000, it_fma = 100000;
  float ms = 0.f;
  RBC_TRY(time_launch([&](uint32_t it) { mfma_bf16_loop<<<grid, kBlock, 0, st>>>(it, out.get()); }, it_mfma, st,
                      &ms));
  *bf16_mfma_tflops = waves * it_mfma * kMfmaChains * (16.0 * 16 * 32 * 2) / (ms * 1e-3) / 1e12;
  RBC_TRY(time_launch([&](uint32_t it) { fma_f32_loop<<<grid, kBlock, 0, st>>>(it, out.get()); }, it_fma, st, &ms));
  *fp32_fma_tflops = waves * 64.0 * it_fma * kFmaChains * 2.0 / (ms * 1e-3) / 1e12;
  return RBC_OK;
}
