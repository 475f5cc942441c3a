// common.h — shared host-side plumbing of libapgx: error reporting across the
// C-ABI and HIP call checking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "apgx.h"

namespace apgx {

// Records `detail` as the calling thread's last error and returns `code`.
int Fail(int code, const std::string& detail);
void ClearError();

// Maps a HIP failure the way the reference maps CUDA failures
// (gpu/cuda/planar_router.cu:67-78): allocation failure is resource
// exhaustion, anything else is a backend failure "<op> failed: <message>".
int HipFail(hipError_t error, const char* operation);

}  // namespace apgx

#define APGX_HIP(expr, op)                              \
  do {                                                  \
    hipError_t apgx_hip_status_ = (expr);               \
    if (apgx_hip_status_ != hipSuccess) {               \
      return ::apgx::HipFail(apgx_hip_status_, (op));   \
    }                                                   \
  } while (0)
