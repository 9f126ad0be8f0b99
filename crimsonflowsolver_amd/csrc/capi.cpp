// Error reporting and device queries of the C ABI.
#include <hip/hip_runtime.h>

#include <cstring>

#include "common.hpp"

namespace crimson {
namespace {
thread_local std::string g_last;
}
void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last = buf;
}
const std::string &last_error() { return g_last; }
}  // namespace crimson

extern "C" int crimson_last_error(char *buf, int len) {
  if (!buf || len <= 0) return CRIMSON_EARG;
  const std::string &s = crimson::last_error();
  const size_t n = std::min((size_t)len - 1, s.size());
  std::memcpy(buf, s.data(), n);
  buf[n] = '\0';
  return CRIMSON_OK;
}

extern "C" int crimson_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
