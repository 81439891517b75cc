// Library-wide C-ABI entry points: error messages and version.
#include <cstring>
#include <string>

#include "common.hpp"

namespace surfhip {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace surfhip

extern "C" size_t surfhip_last_error(char *buf, size_t n) {
  const std::string &s = surfhip::g_last_error;
  if (buf && n) {
    size_t m = s.size() < n - 1 ? s.size() : n - 1;
    std::memcpy(buf, s.data(), m);
    buf[m] = '\0';
  }
  return s.size();
}

extern "C" const char *surfhip_version(void) { return "surfhip 0.1 gfx950"; }
