// Error reporting and device checks shared by the C-ABI translation units of libcloudsc_hip.so.
#pragma once

#include <hip/hip_runtime.h>

namespace cloudsc_capi {
// Records `what` (and the HIP error text, if any) as the calling thread's cloudsc_last_error() and returns code.
int fail(int code, const char* what, hipError_t e = hipSuccess);
// CLOUDSC_OK if the current device is a gfx950, CLOUDSC_ENODEVICE otherwise (result cached per device).
int check_device();
}  // namespace cloudsc_capi
