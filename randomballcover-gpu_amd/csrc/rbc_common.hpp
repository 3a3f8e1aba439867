// Shared host-side helpers of librbc_hip: status codes, the per-thread error message,
// HIP error checking and a small owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>

#include "rbc_hip.h"

namespace rbc {

void set_error(const char *fmt, ...);
// RBC_ERR_CHECK (with the failing unit and line as the message) if a device bounds check
// failed since the last call; RBC_OK otherwise and always in builds without RBC_CHECK.
int device_check();

#define RBC_HIP_TRY(expr)                                                                     \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      ::rbc::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,                    \
                       hipGetErrorString(_e));                                                \
      return _e == hipErrorOutOfMemory ? RBC_ERR_NOMEM : RBC_ERR_HIP;                         \
    }                                                                                         \
  } while (0)

#define RBC_TRY(expr)                                                                         \
  do {                                                                                        \
    int _s = (expr);                                                                          \
    if (_s != RBC_OK) return _s;                                                              \
  } while (0)

#define RBC_ARG_CHECK(cond, ...)                                                              \
  do {                                                                                        \
    if (!(cond)) {                                                                            \
      ::rbc::set_error(__VA_ARGS__);                                                          \
      return RBC_ERR_ARG;                                                                     \
    }                                                                                         \
  } while (0)

// Owning device allocation.  Move-only; freed on destruction.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    reset();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? RBC_ERR_NOMEM : RBC_ERR_HIP;
    }
    n = count;
    return RBC_OK;
  }
  // Grow-only reuse: keeps the allocation when it is already large enough.
  int ensure(size_t count) { return count <= n && p ? RBC_OK : alloc(count); }
  T *get() const { return p; }
};

inline uint32_t pad16(uint32_t v) { return (v + 15u) & ~15u; }
inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
inline size_t cdivz(size_t a, size_t b) { return (a + b - 1) / b; }

// Returns RBC_OK when the pointer refers to device-accessible memory.
int check_device_ptr(const void *p, const char *what);

}  // namespace rbc
