// Host-side error reporting of libwt: a thread-local message buffer and the
// status-returning fail(). Free of HIP, so host-only units (the voxelizer) and
// their sanitizer builds compile it with any C++ compiler.
#pragma once

#include <cstdarg>
#include <cstdio>

namespace wt {

char* error_buffer();  // thread-local, 512 bytes

// Record a formatted error message and return `code`.
inline int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(error_buffer(), 512, fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace wt
