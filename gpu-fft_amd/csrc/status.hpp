// Status codes and per-thread error detail of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/gpufft_amd.h"

namespace gfa {

// Records a message for gpufft_last_error_message() and returns `status`.
int fail(int status, const std::string& msg);
// Records a HIP error (hipGetLastError after a launch) as GPUFFT_ERR_HIP.
int fail_hip(hipError_t e, const char* where);
// Clears the per-thread error state; returns GPUFFT_OK.
int ok();

}  // namespace gfa

#define GFA_CHECK_LAUNCH(where)                                  \
    do {                                                         \
        hipError_t e_ = hipGetLastError();                       \
        if (e_ != hipSuccess) return ::gfa::fail_hip(e_, where); \
    } while (0)
