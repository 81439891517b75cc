// Shared helpers of libsurfhip: status/error plumbing and small utilities.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/surfhip.h"

#if defined(__HIPCC__)
#define SH_HD __host__ __device__
#else
#define SH_HD
#endif

namespace surfhip {

// An error carrying the C-ABI status code; thrown by host code and turned
// into a status + thread-local message at the C boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string &msg);

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

inline std::string fmt(const char *f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// Run `body`, translating exceptions into a status code and message.
template <class F> int guarded(F &&body) {
  try {
    body();
    return SURFHIP_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return SURFHIP_ERR_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return SURFHIP_ERR_DEVICE;
  }
}

}  // namespace surfhip
