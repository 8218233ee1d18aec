// Error reporting and device checks shared by the C-ABI translation units of libcloudsc_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cloudsc_hip.h"

namespace cloudsc_capi {
// Records `what` (and the HIP error text, if any) as the calling thread's cloudsc_last_error() and returns code.
int fail(int code, const char* what, hipError_t e = hipSuccess);
// CLOUDSC_OK if the current device is a gfx950, CLOUDSC_ENODEVICE otherwise (result cached per device).
int check_device();
// The switches of the parameter block that the kernels implement: CLOUDSC_EUNSUPPORTED for LCLOUD_INHOMOG,
// LRAD_CLOUD_INHOMOG, aerosol, SPP, unknown lcldbud bits and, unless `budget` (the entry takes PEXTRA), the
// budget switches that write PEXTRA.
int check_config(const cloudsc_params_t* p, bool budget = false);
}  // namespace cloudsc_capi
