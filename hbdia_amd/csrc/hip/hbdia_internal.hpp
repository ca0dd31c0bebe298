// hbdia_internal.hpp -- helpers shared by the HIP translation units of
// libhbdia_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace hbdia_detail {

// The thread's last error, returned by hbdia_last_error().
extern thread_local std::string g_last_error;
// Records msg as the last error and returns code.
int fail(int code, const std::string& msg);
// HBDIA_SUCCESS for hipSuccess, else records "what: <HIP error>" and returns
// HBDIA_ERROR_OUT_OF_MEMORY or HBDIA_ERROR_DEVICE.
int check_hip(hipError_t e, const char* what);

}  // namespace hbdia_detail
