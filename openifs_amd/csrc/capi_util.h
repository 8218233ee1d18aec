// Error reporting and device checks shared by the C-ABI translation units of libcloudsc_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cloudsc_hip.h"

namespace cloudsc_capi {
// Records `what` (and the HIP error text, if any) as the calling thread's cloudsc_last_error() and returns code.
int fail(int code, const char* what, hipError_t e = hipSuccess);
// CLOUDSC_OK if the current device is a gfx950, CLOUDSC_ENODEVICE otherwise (result cached per device).
int check_device();
// The switches of the parameter block that the kernels implement: CLOUDSC_EUNSUPPORTED for aerosol, SPP, unknown
// lcldbud bits and, unless `diag` (the entry takes PEXTRA, PGAW and PFSD), the budget switches that write PEXTRA,
// LRAD_CLOUD_INHOMOG, which writes PFSD, and LCLOUD_INHOMOG, which reads PGAW.
int check_config(const cloudsc_params_t* p, bool diag = false);
// CLOUDSC_ENULL if LRAD_CLOUD_INHOMOG is on and PGAW or PFSD is NULL, or LCLOUD_INHOMOG is on and PGAW is NULL.
int check_fsd(const cloudsc_params_t* p, const void* pgaw, const void* pfsd);

// Device buffer and stream of the per-block host entries.  Every such entry copies its host arrays in, launches,
// copies back and waits on this stream before it returns, so the entries of one thread share the calling
// thread's workspace, and threads never share one (reentrancy under the OpenMP block loop).
struct HostWorkspace {
  void* dev = nullptr;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  int device = -1;
};
// The calling thread's workspace on the current device, grown to at least `bytes`.  When the thread has moved to
// another device, the stream and buffer held on the previous one are released under that device first; the new
// device is recorded only once its stream exists.  `who` prefixes the error text.
int thread_workspace(size_t bytes, const char* who, HostWorkspace** out);
}  // namespace cloudsc_capi
