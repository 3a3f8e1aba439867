// Library-wide state and helpers: per-thread error text, device selection, metric
// setting, input staging and small fill kernels.
#include <cstring>
#include <mutex>

#include <vector>

#include "launch.hpp"

namespace rbc {

namespace {
thread_local char g_err[1024] = "";
int g_metric = RBC_L2;

__global__ void fill_u32_k(uint32_t *p, uint32_t v, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void fill_f32_k(float *p, float v, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
}  // namespace

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int metric_setting() { return g_metric; }

namespace {
struct CheckReader {
  int (*read)(unsigned int *);
  const char *unit;
};
std::vector<CheckReader> &check_readers() {
  static std::vector<CheckReader> r;
  return r;
}
}  // namespace

void register_check_reader(int (*read)(unsigned int *), const char *unit) {
  check_readers().push_back(CheckReader{read, unit});
}

int device_check() {
  for (const CheckReader &r : check_readers()) {
    unsigned int line = 0;
    if (r.read(&line) != 0) {
      set_error("device bounds check: reading the check word of %s failed", r.unit);
      return RBC_ERR_HIP;
    }
    if (line) {
      set_error("device bounds check failed in %s, line %u (of that file or a header it includes)", r.unit, line);
      return RBC_ERR_CHECK;
    }
  }
  return RBC_OK;
}

int max_lds_bytes() {
  static int v = 0;
  if (!v) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 65536;
    v = (int)prop.sharedMemPerBlock;
  }
  return v;
}

int check_device_ptr(const void *p, const char *what) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    set_error("%s is not a device pointer", what);
    return RBC_ERR_ARG;
  }
  return RBC_OK;
}

int fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st) {
  if (!n) return RBC_OK;
  uint32_t g = (uint32_t)std::min<size_t>(cdivz(n, 256), 4096);
  hipLaunchKernelGGL(fill_u32_k, dim3(g), dim3(256), 0, st, p, v, n);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int fill_f32(float *p, float v, size_t n, hipStream_t st) {
  if (!n) return RBC_OK;
  uint32_t g = (uint32_t)std::min<size_t>(cdivz(n, 256), 4096);
  hipLaunchKernelGGL(fill_f32_k, dim3(g), dim3(256), 0, st, p, v, n);
  RBC_HIP_TRY(hipGetLastError());
  return RBC_OK;
}

int DevMat::stage(const float *x, uint32_t rows, uint32_t d, uint32_t ldx, int mem, hipStream_t st) {
  RBC_ARG_CHECK(x != nullptr || rows == 0, "matrix pointer is NULL");
  RBC_ARG_CHECK(ldx >= d, "leading dimension %u < dimension %u", ldx, d);
  if (mem == RBC_MEM_DEVICE && ldx % 4 == 0) {
    p = x;
    ld = ldx;
    return RBC_OK;
  }
  ld = pad16(d);
  RBC_TRY(own.alloc(size_t(rows) * ld));
  if (ld != d) RBC_TRY(fill_f32(own.get(), 0.f, size_t(rows) * ld, st));
  if (rows)
    RBC_TRY(rows_in(own.get(), ld * sizeof(float), x, size_t(ldx) * sizeof(float), size_t(d) * sizeof(float),
                    rows, mem, st));
  p = own.get();
  return RBC_OK;
}

int rows_in(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows, int mem,
            hipStream_t st) {
  if (!rows || !width) return RBC_OK;
  const size_t span = (rows - 1) * spitch + width;  // bytes the source rows cover
  if (mem == RBC_MEM_DEVICE) {
    RBC_HIP_TRY(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToDevice, st));
    return RBC_OK;
  }
  if (spitch == dpitch || rows == 1) {
    RBC_HIP_TRY(hipMemcpyAsync(dst, src, span, hipMemcpyHostToDevice, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
    return RBC_OK;
  }
  DevBuf<unsigned char> tmp;
  RBC_TRY(tmp.alloc(span));
  RBC_HIP_TRY(hipMemcpyAsync(tmp.get(), src, span, hipMemcpyHostToDevice, st));
  RBC_HIP_TRY(hipMemcpy2DAsync(dst, dpitch, tmp.get(), spitch, width, rows, hipMemcpyDeviceToDevice, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int rows_out(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows, int mem,
             hipStream_t st) {
  if (!rows || !width) return RBC_OK;
  if (mem == RBC_MEM_DEVICE) {
    RBC_HIP_TRY(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToDevice, st));
    return RBC_OK;
  }
  if (spitch == dpitch || rows == 1) {
    RBC_HIP_TRY(hipMemcpyAsync(dst, src, (rows - 1) * spitch + width, hipMemcpyDeviceToHost, st));
    RBC_HIP_TRY(hipStreamSynchronize(st));
    return RBC_OK;
  }
  std::vector<unsigned char> tmp((rows - 1) * spitch + width);
  RBC_HIP_TRY(hipMemcpyAsync(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  unsigned char *o = static_cast<unsigned char *>(dst);
  for (size_t r = 0; r < rows; ++r) std::memcpy(o + r * dpitch, tmp.data() + r * spitch, width);
  return RBC_OK;
}

int upload_host(void *dst, const void *src, size_t bytes, hipStream_t st) {
  if (!bytes) return RBC_OK;
  RBC_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int copy_out(void *dst, const void *src, size_t bytes, int mem, hipStream_t st) {
  if (!bytes) return RBC_OK;
  RBC_HIP_TRY(hipMemcpyAsync(dst, src, bytes,
                             mem == RBC_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                             st));
  // The caller reads a host destination as soon as the call returns.
  if (mem != RBC_MEM_DEVICE) RBC_HIP_TRY(hipStreamSynchronize(st));
  return RBC_OK;
}

int copy_in(void *dst, const void *src, size_t bytes, int mem, hipStream_t st) {
  if (!bytes) return RBC_OK;
  if (mem != RBC_MEM_DEVICE) return upload_host(dst, src, bytes, st);
  RBC_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
  return RBC_OK;
}

int upload_items(const std::vector<ScanItem> &v, DevBuf<ScanItem> &buf, hipStream_t st) {
  RBC_TRY(buf.ensure(v.size()));
  return upload_host(buf.get(), v.data(), v.size() * sizeof(ScanItem), st);
}

}  // namespace rbc

extern "C" {

const char *rbc_last_error(void) { return rbc::g_err; }

int rbc_version(void) { return 100; }

int rbc_set_device(int device) {
  RBC_HIP_TRY(hipSetDevice(device));
  return RBC_OK;
}

int rbc_set_metric(int metric) {
  RBC_ARG_CHECK(metric == RBC_L2 || metric == RBC_L1, "unknown metric %d", metric);
  rbc::g_metric = metric;
  return RBC_OK;
}

int rbc_get_metric(void) { return rbc::g_metric; }

}  // extern "C"

namespace {
__global__ void check_probe(const uint32_t *buf, uint32_t index, uint32_t bound, uint32_t *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = buf[RBC_CHECKED(index, bound)];
}
}  // namespace

extern "C" int rbc_check_enabled(void) {
#ifdef RBC_CHECK
  return 1;
#else
  return 0;
#endif
}

extern "C" int rbc_check_selftest(uint32_t index, uint32_t bound, uint32_t *value) {
  RBC_ARG_CHECK(bound >= 1 && bound <= (1u << 20) && value, "bound %u outside [1, 2^20] or NULL value", bound);
  // Without the checks an index past the buffer would be a real out-of-bounds read.
  RBC_ARG_CHECK(rbc_check_enabled() || index < bound, "index %u >= bound %u needs a build with RBC_CHECK", index,
                bound);
  using namespace rbc;
  std::vector<uint32_t> h(bound);
  for (uint32_t i = 0; i < bound; ++i) h[i] = 1000u + i;
  DevBuf<uint32_t> buf, out;
  RBC_TRY(buf.alloc(bound));
  RBC_TRY(out.alloc(1));
  RBC_TRY(upload_host(buf.get(), h.data(), size_t(bound) * 4, nullptr));
  hipLaunchKernelGGL(check_probe, dim3(1), dim3(64), 0, nullptr, buf.get(), index, bound, out.get());
  RBC_HIP_TRY(hipGetLastError());
  RBC_HIP_TRY(hipMemcpy(value, out.get(), 4, hipMemcpyDeviceToHost));
  return device_check();
}
