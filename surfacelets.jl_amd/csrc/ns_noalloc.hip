// surfhip_ns_analyze_noalloc: analyze! that allocates no device memory.
//
// For callers that cannot allocate (code under HIP graph capture, tight
// serving loops): the plan's analysis runs with its working spectra and
// slabs carved from `scratch`, a file-scope device array of fixed size, not
// from buffers the library allocates.  The plan itself (its tables) is built
// beforehand by surfhip_ns_plan_create.
//
// `scratch` is a host-side stand-in in host code; the device address of
// each GPU's copy comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address is passed to kernels.  Every
// call on one GPU shares that GPU's array: calls that overlap in time on one
// GPU (on different streams, or from different threads) race on it.  The
// analysis writes every scratch element it later reads (the slabs' dead
// tiles are read only to be discarded by the liveness masks), so nothing in
// it needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hip_util.hpp"
#include "ns_scratch.hpp"

namespace surfhip {
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// Device address of `dev`'s copy of scratch, looked up once per device with
// that device current (and the caller's current device restored).
void *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) fail(SURFHIP_ERR_DEVICE, fmt("device %d outside 0:%d", dev, kMaxDevices - 1));
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[dev]) return g_ptr[dev];
  int cur = 0;
  SH_HIP(hipGetDevice(&cur));
  if (cur != dev) SH_HIP(hipSetDevice(dev));
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || p == nullptr)
    fail(SURFHIP_ERR_DEVICE, fmt("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e)));
  g_ptr[dev] = p;
  return p;
}

}  // namespace
}  // namespace surfhip

using namespace surfhip;

extern "C" int surfhip_ns_analyze_noalloc(surfhip_ns_plan *plan, const void *x, void *coeffs) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    const size_t need = ns_analyze_working_bytes(plan);
    if (need > kScratchBytes)
      fail(SURFHIP_ERR_ARG, fmt("this plan's analysis needs %zu bytes of working memory; the allocation-free "
                                "entry point's static scratch holds %zu (use surfhip_ns_analyze)",
                                need, kScratchBytes));
    ns_analyze_in(plan, scratch_on(ns_plan_device(plan)), kScratchBytes, x, coeffs);
  });
}
