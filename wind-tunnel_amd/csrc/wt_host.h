// Host-side helpers of libwt: thread-local error text and launch checks.
#pragma once

#include <hip/hip_runtime.h>

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

// Turn the last launch error into a status.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(-2, "%s: launch failed: %s", what, hipGetErrorString(e));
  return 0;
}

}  // namespace wt
