// Host-side helpers of libwt: error reporting (wt_error.h) and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include "wt_error.h"

namespace wt {

// Turn the last launch error into a status.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(-2, "%s: launch failed: %s", what, hipGetErrorString(e));
  return 0;
}

}  // namespace wt
