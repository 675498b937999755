// Host-side launchers of the kernels (one per .hip file); called by the C ABI
// in tk_api.hip after argument validation.  Each returns 0 or -1 for an
// unsupported form; HIP launch errors are read by the caller.
#pragma once

#include <hip/hip_runtime.h>

#include "tk_internal.h"

namespace tk {

int launch_butterfly(bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                     const DeviceTables *tab, int grid_cap, hipStream_t stream);

int launch_mfma(int shape, bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                const DeviceTables *tab, int grid_cap, int waves, hipStream_t stream);

int launch_basemul(int form, int16_t *r, const int16_t *a, const int16_t *b, size_t npoly,
                   const DeviceTables *tab, int grid_cap, hipStream_t stream);

}  // namespace tk
