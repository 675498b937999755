// Whole-poly product in the NTT domain: ref basemul (ref/ntt.c:139-146) on
// each of the 128 degree-1 pairs, pair (4i, 4i+1) with zetas[64+i] and pair
// (4i+2, 4i+3) with -zetas[64+i].  Streaming, HBM-bound (3 x 512 B per
// poly); each lane handles 8 consecutive coefficients with 16-byte accesses.
//
// The grid covers the batch, one 16-byte chunk per lane: on MI355X that
// streams at 6.2 TB/s, where a grid of the resident workgroups looping over
// the batch reaches 4.5 TB/s (4.9-5.1 with 2-4 chunks per lane in flight).
// The loop only takes over beyond the largest grid (2^32 - 1 lanes: 2^27
// polys).
#include "tk_device.h"
#include "tk_grid.h"
#include "tk_launch.h"

namespace tk {
namespace {

__device__ __forceinline__ void pair(int16_t &r0, int16_t &r1, int16_t a0, int16_t a1, int16_t b0,
                                     int16_t b1, int16_t z)
{
  int16_t t = fqmul(a1, b1);
  t = fqmul(t, z);
  r0 = (int16_t)(t + fqmul(a0, b0));
  int16_t u = fqmul(a0, b1);
  r1 = (int16_t)(u + fqmul(a1, b0));
}

template <int FORM>
__global__ __launch_bounds__(256) void k_basemul(int16_t *__restrict__ r,
                                                 const int16_t *__restrict__ a,
                                                 const int16_t *__restrict__ b, size_t npoly,
                                                 const DeviceTables *__restrict__ tab)
{
  const size_t nchunk = npoly * (kN / 8);
  for (size_t ch = (size_t)blockIdx.x * blockDim.x + threadIdx.x; ch < nchunk;
       ch += (size_t)gridDim.x * blockDim.x) {
    union U {
      v4i v;
      int16_t h[8];
    } va, vb, vr;
    va.v = ld_stream(reinterpret_cast<const v4i *>(a) + ch);
    vb.v = ld_stream(reinterpret_cast<const v4i *>(b) + ch);
    const int i0 = (int)(ch & 31) * 2;  // group of 4 coefficients within the poly
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int16_t z = tab->zetas[64 + i0 + q];
      pair(vr.h[4 * q], vr.h[4 * q + 1], va.h[4 * q], va.h[4 * q + 1], vb.h[4 * q],
           vb.h[4 * q + 1], z);
      pair(vr.h[4 * q + 2], vr.h[4 * q + 3], va.h[4 * q + 2], va.h[4 * q + 3], vb.h[4 * q + 2],
           vb.h[4 * q + 3], (int16_t)-z);
    }
    if (FORM == kFormCanon) {
#pragma unroll
      for (int j = 0; j < 8; j++)
        vr.h[j] = canon16(vr.h[j]);
    } else if (FORM == kFormCentred) {
#pragma unroll
      for (int j = 0; j < 8; j++)
        vr.h[j] = centre16(vr.h[j]);
    }
    st_stream(reinterpret_cast<v4i *>(r) + ch, vr.v);
  }
}

}  // namespace

int launch_basemul(int form, int16_t *r, const int16_t *a, const int16_t *b, size_t npoly,
                   const DeviceTables *tab, int ncu, hipStream_t stream)
{
  (void)ncu;
  const unsigned grid = stream_grid(npoly * (kN / 8), 256);
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, stream, r, a, b, npoly, tab);
    return 0;
  };
  switch (form) {
    case kFormRef: return go(k_basemul<kFormRef>);
    case kFormCanon: return go(k_basemul<kFormCanon>);
    case kFormCentred: return go(k_basemul<kFormCentred>);
    default: return -1;
  }
}

}  // namespace tk
