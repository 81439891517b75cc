// HIP error checking and small device buffers for libsurfhip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
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
    // A request larger than the whole device can never succeed, and one
    // above SURFHIP_MAX_ALLOCATION (bytes; for processes that must fail
    // rather than oversubscribe device memory, which the runtime allows
    // instead of failing an allocation) is refused: both are reported as
    // allocation failures without asking the runtime.
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0 && count > total_b / sizeof(T))
      fail(SURFHIP_ERR_OOM, fmt("device allocation of %zu bytes exceeds the device's %zu bytes",
                                count * sizeof(T), total_b));
    if (const char *cap = std::getenv("SURFHIP_MAX_ALLOCATION")) {
      const unsigned long long lim = std::strtoull(cap, nullptr, 10);
      if (lim > 0 && count > lim / sizeof(T))
        fail(SURFHIP_ERR_OOM, fmt("device allocation of %zu bytes exceeds SURFHIP_MAX_ALLOCATION = %llu bytes",
                                  count * sizeof(T), lim));
    }
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

// Check, on the host and before anything is launched, that `bytes` bytes
// from `ptr` lie in one device allocation (hipMemGetAddressRange); a buffer
// that ends early fails with SURFHIP_ERR_DIM (the reference's
// DimensionMismatch), memory the device cannot address with SURFHIP_ERR_ARG.
// Pinned or managed memory, whose extent the runtime does not report this
// way, is accepted unchecked.  A sub-allocation of a larger block (a caching
// allocator's) is checked against the block it lies in.
inline void check_device_buffer(const void *ptr, size_t bytes, const char *what) {
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)ptr) == hipSuccess && base != nullptr) {
    const size_t have = (size_t)((const char *)base + size - (const char *)ptr);
    if ((const char *)ptr < (const char *)base || have < bytes)
      fail(SURFHIP_ERR_DIM, fmt("the %s holds %zu bytes from the pointer passed, the plan needs %zu", what,
                                (const char *)ptr < (const char *)base ? (size_t)0 : have, bytes));
    return;
  }
  (void)hipGetLastError();  // the failed lookup must not be reported by a later launch check
  hipPointerAttribute_t a{};
  const bool known = hipPointerGetAttributes(&a, ptr) == hipSuccess;
  (void)hipGetLastError();
  const bool device_visible =
      known && (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged ||
                (a.type == hipMemoryTypeHost && a.devicePointer != nullptr));
  if (!device_visible) fail(SURFHIP_ERR_ARG, fmt("the %s is not memory the device can address", what));
}

}  // namespace surfhip
