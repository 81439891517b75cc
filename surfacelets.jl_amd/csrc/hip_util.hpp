// HIP error checking and small device buffers for libsurfhip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

#define SH_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      ::surfhip::fail(e_ == hipErrorOutOfMemory ? SURFHIP_ERR_OOM : SURFHIP_ERR_DEVICE,     \
                      ::surfhip::fmt("%s failed: %s", #call, hipGetErrorString(e_)));       \
  } while (0)

#define SH_LAUNCHED() SH_HIP(hipGetLastError())

namespace surfhip {

// Owning device allocation.
template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      fail(e == hipErrorOutOfMemory ? SURFHIP_ERR_OOM : SURFHIP_ERR_DEVICE,
           fmt("device allocation of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e)));
    }
    n = count;
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size());
    if (!v.empty()) SH_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
};

}  // namespace surfhip
