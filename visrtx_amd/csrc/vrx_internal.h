// vrx_internal.h -- error plumbing shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../../include/vrx_hip.h"

namespace vrx {
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
// compute units of the current device (cached per device; 256 if unknown)
int device_cus();
} // namespace vrx

#define VRX_HIP(call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess)                                               \
      return vrx::hip_fail(e_, #call, __FILE__, __LINE__);              \
  } while (0)

#define VRX_LAUNCH_CHECK()                                              \
  do {                                                                  \
    hipError_t e_ = hipGetLastError();                                  \
    if (e_ != hipSuccess)                                               \
      return vrx::hip_fail(e_, "kernel launch", __FILE__, __LINE__);    \
  } while (0)

#define VRX_ARG(cond, msg)                                              \
  do {                                                                  \
    if (!(cond)) {                                                      \
      vrx::set_error("invalid argument: %s", msg);                      \
      return VRX_ERR_ARG;                                               \
    }                                                                   \
  } while (0)
