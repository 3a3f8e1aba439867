// Host-side launch helpers for the tile engine and the input staging shared by the
// entry points.
#pragma once

#include <vector>

#include "prims.hpp"
#include "rbc_tile.hpp"

namespace rbc {

int max_lds_bytes();

template <int METRIC, class Epi>
int launch_tiles(const TileSrc &src, const ScanItem *d_items, uint32_t nitems, const Epi &epi,
                 hipStream_t st) {
  if (nitems == 0) return RBC_OK;
  size_t lds = tile_lds_base_bytes(src.dpad) + Epi::lds_bytes();
  if ((int)lds > max_lds_bytes()) {
    set_error("tile kernel needs %zu bytes of LDS (d=%u); the device has %d", lds, src.d,
              max_lds_bytes());
    return RBC_ERR_LIMIT;
  }
  static bool attr_set = false;
  if (!attr_set) {
    RBC_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&tile_kernel<METRIC, Epi>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes()));
    attr_set = true;
  }
  hipLaunchKernelGGL((tile_kernel<METRIC, Epi>), dim3(nitems), dim3(kNT), lds, st, src, d_items, epi);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

// Dispatch on the runtime metric.
template <class Epi>
int launch_tiles_m(int metric, const TileSrc &src, const ScanItem *d_items, uint32_t nitems,
                   const Epi &epi, hipStream_t st) {
  return metric == RBC_L1 ? launch_tiles<1>(src, d_items, nitems, epi, st)
                          : launch_tiles<0>(src, d_items, nitems, epi, st);
}

// A float matrix on the device with a leading dimension that is a multiple of 4.
// Either borrows the caller's device pointer or owns a padded copy.
struct DevMat {
  const float *p = nullptr;
  size_t ld = 0;
  DevBuf<float> own;
  // Stages rows x (host or device, leading dimension ldx) of width d.
  int stage(const float *x, uint32_t rows, uint32_t d, uint32_t ldx, int mem, hipStream_t st);
};

// Contiguous copies between the device and the caller's buffer (host or device).  With a
// host buffer both return after the copy is done, so the buffer may be reused (copy_in)
// or read (copy_out) at once; device-to-device copies stay asynchronous on st.
int copy_out(void *dst, const void *src, size_t bytes, int mem, hipStream_t st);
int copy_in(void *dst, const void *src, size_t bytes, int mem, hipStream_t st);
// Host-to-device copy that returns once the host buffer may be modified or freed.
int upload_host(void *dst, const void *src, size_t bytes, hipStream_t st);

// Row-strided copies: `rows` rows of `width` bytes, source pitch `spitch`, destination
// pitch `dpitch` (bytes).  The host end of a transfer is always moved as one contiguous
// block (a strided host transfer can degrade to one DMA per row); re-pitching happens on
// the device (rows_in) or by host memcpy (rows_out).  Both return with the copy done when
// the caller's end is host memory.
int rows_in(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows, int mem,
            hipStream_t st);
int rows_out(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows, int mem,
             hipStream_t st);

// Uploads a host vector of items.
int upload_items(const std::vector<ScanItem> &v, DevBuf<ScanItem> &buf, hipStream_t st);

// Kernel-side fill (graph-capture friendly replacement for hipMemsetAsync).
int fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);
int fill_f32(float *p, float v, size_t n, hipStream_t st);

}  // namespace rbc
