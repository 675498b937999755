// Chosen method: the seven NTT layers as two exact INT8 matrix products on
// the CDNA4 matrix cores.
//
// A poly is x[m][c] = r[16m + c].  Forward: layers 1-4 are one 16x16 matrix
// over m (the same for every c), layers 5-7 are one 16x16 matrix per block m
// over c.  Inverse: the per-block matrices first, then one matrix over m that
// also carries the final 1441 scaling (ref/ntt.c:124-125).  Every constant is
// split into signed 8-bit limbs, the data into two signed byte digits (see
// DeviceTables in tk_internal.h), the products accumulate exactly in int32,
// and each output is reduced once with a 24-bit Barrett step.  The result is
// exact mod q for every int16 input.
//
// Data is the MFMA B operand (one poly per output column, so a wave works on
// 16 or 32 polys at once) and the constants are the A operand, resident in
// VGPRs for the life of the wave.  Each wave owns an LDS slab holding its
// tile; the two passes read and write only lane-private LDS words, so the only
// cross-lane hand-offs are tile load -> pass 1 -> pass 2 -> tile store.
#include "tk_launch.h"
#include "tk_mfma_core.h"

namespace tk {
namespace {

// ---------------------------------------------------------------------------
// Kernels.  INV selects the inverse; FORM the output form of the last pass.
// Grid-stride over tiles of P polys per wave; WAVES waves per workgroup.
// ---------------------------------------------------------------------------
template <int FORM, bool INV>
__device__ __forceinline__ int32_t bias_adjust()
{
  // Forward tables carry no centring offset, inverse final tables do.
  if (!INV)
    return FORM == kFormCentred ? kCentre : 0;
  return FORM == kFormCentred ? 0 : -kCentre;
}

// Per-tile poly addresses.  Linear batches compute them; ragged batches walk
// the request table once per tile (lanes 0..P-1, one poly each) and park the
// pointers in LDS for the staging loops.
template <int P>
struct TileAddr {
  const int16_t *in0;
  int16_t *out0;
  int16_t **lds;  // ragged only
  size_t valid;   // linear only: polys of this tile inside the batch
  __device__ const int16_t *src(int pp) const
  {
    if (lds)
      return lds[pp];
    return (size_t)pp < valid ? in0 + (size_t)pp * kN : nullptr;
  }
  __device__ int16_t *dst(int pp) const
  {
    if (lds)
      return lds[pp];
    return (size_t)pp < valid ? out0 + (size_t)pp * kN : nullptr;
  }
};

template <int P>
__device__ __forceinline__ TileAddr<P> tile_addr(const LinearTiles &io, size_t tile, int lane,
                                                 int16_t **lds)
{
  (void)lane;
  (void)lds;
  const size_t p0 = tile * P;
  const size_t left = io.npoly - p0;
  return TileAddr<P>{io.in + p0 * kN, io.out + p0 * kN, nullptr, left < (size_t)P ? left : P};
}

template <int P>
__device__ __forceinline__ TileAddr<P> tile_addr(const RaggedTiles &io, size_t tile, int lane,
                                                 int16_t **lds)
{
  if (lane < P) {
    const uint64_t g = (uint64_t)tile * P + lane;
    int16_t *p = nullptr;
    if (g < io.npoly) {
      uint32_t r = io.tile_req[tile];
      while (io.first[r + 1] <= g)
        r++;
      p = io.base[r] + (size_t)(g - io.first[r]) * kN;
    }
    lds[lane] = p;
  }
  wave_sync();
  return TileAddr<P>{nullptr, nullptr, lds, 0};
}

template <int FORM, bool INV, int WAVES, class Tiles>
__global__ __launch_bounds__(64 * WAVES) void k_mfma16(Tiles io,
                                                        const DeviceTables *__restrict__ tab)
{
  constexpr int P = 16;
  __shared__ __align__(16) int16_t s_poly[WAVES][P * kStride];
  __shared__ int16_t *s_ptr[WAVES][P];
  __shared__ __align__(16) int32_t s_bias[16][16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, t = lane & 15;

  const int32_t(*bias_c)[16] = INV ? tab->bias_inv2 : tab->bias_fwd2;
  for (int i = threadIdx.x; i < 256; i += 64 * WAVES)
    s_bias[i >> 4][i & 15] = bias_c[i >> 4][i & 15] + (INV ? 0 : bias_adjust<FORM, INV>());
  __syncthreads();

  const auto &m1 = INV ? tab->m16_inv1[0] : tab->m16_fwd1[0];
  const auto &m2 = INV ? tab->m16_inv2 : tab->m16_fwd2;
  const long a1lo = *reinterpret_cast<const long *>(m1[0][lane]);
  const long a1hi = *reinterpret_cast<const long *>(m1[1][lane]);
  const int32_t *b1p = INV ? tab->bias_inv1[0] : tab->bias_fwd1[0];
  const int32_t b1adj = INV ? bias_adjust<FORM, INV>() : 0;
  const v4i bias1 = {b1p[4 * g] + b1adj, b1p[4 * g + 1] + b1adj, b1p[4 * g + 2] + b1adj,
                     b1p[4 * g + 3] + b1adj};
  long a2lo[16], a2hi[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    a2lo[m] = *reinterpret_cast<const long *>(m2[m][0][lane]);
    a2hi[m] = *reinterpret_cast<const long *>(m2[m][1][lane]);
  }

  int16_t *L = s_poly[w];
  int16_t *Pp = L + t * kStride;
  const size_t ntiles = (io.npoly + P - 1) / P;
  for (size_t tile = (size_t)blockIdx.x * WAVES + w; tile < ntiles;
       tile += (size_t)gridDim.x * WAVES) {
    const TileAddr<P> ta = tile_addr<P>(io, tile, lane, s_ptr[w]);
    tile_in<P>(L, [&](int pp) { return ta.src(pp); }, lane);
    wave_sync();
    if (!INV) {
      pass_m16<kFormInter>(Pp, g, a1lo, a1hi, bias1);
      wave_sync();
      pass_c16<FORM>(Pp, g, [&](int m) { return a2lo[m]; }, [&](int m) { return a2hi[m]; },
                     s_bias);
    } else {
      pass_c16<kFormInter>(Pp, g, [&](int m) { return a2lo[m]; },
                           [&](int m) { return a2hi[m]; }, s_bias);
      wave_sync();
      pass_m16<FORM>(Pp, g, a1lo, a1hi, bias1);
    }
    wave_sync();
    tile_out<P>(L, [&](int pp) { return ta.dst(pp); }, lane);
    wave_sync();
  }
}

template <int FORM, bool INV, int WAVES, class Tiles>
__global__ __launch_bounds__(64 * WAVES) void k_mfma32(Tiles io,
                                                        const DeviceTables *__restrict__ tab)
{
  constexpr int P = 32;
  __shared__ __align__(16) int16_t s_poly[WAVES][P * kStride];
  __shared__ int16_t *s_ptr[WAVES][P];
  __shared__ __align__(16) int32_t s_bias[16][16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, t = lane & 31;

  const int32_t(*bias_c)[16] = INV ? tab->bias_inv2 : tab->bias_fwd2;
  for (int i = threadIdx.x; i < 256; i += 64 * WAVES)
    s_bias[i >> 4][i & 15] = bias_c[i >> 4][i & 15] + (INV ? 0 : bias_adjust<FORM, INV>());
  __syncthreads();

  const auto &m1 = INV ? tab->m32_inv1[0] : tab->m32_fwd1[0];
  const auto &m2 = INV ? tab->m32_inv2 : tab->m32_fwd2;
  const v4i a1 = *reinterpret_cast<const v4i *>(m1[lane]);
  const int32_t *b1p = INV ? tab->bias_inv1[0] : tab->bias_fwd1[0];
  const int32_t b1adj = INV ? bias_adjust<FORM, INV>() : 0;
  const v4i bias1a = {b1p[4 * h] + b1adj, b1p[4 * h + 1] + b1adj, b1p[4 * h + 2] + b1adj,
                      b1p[4 * h + 3] + b1adj};
  const v4i bias1b = {b1p[8 + 4 * h] + b1adj, b1p[8 + 4 * h + 1] + b1adj,
                      b1p[8 + 4 * h + 2] + b1adj, b1p[8 + 4 * h + 3] + b1adj};
  v4i a2[16];
#pragma unroll
  for (int m = 0; m < 16; m++)
    a2[m] = *reinterpret_cast<const v4i *>(m2[m][lane]);

  int16_t *L = s_poly[w];
  int16_t *Pp = L + t * kStride;
  const size_t ntiles = (io.npoly + P - 1) / P;
  for (size_t tile = (size_t)blockIdx.x * WAVES + w; tile < ntiles;
       tile += (size_t)gridDim.x * WAVES) {
    const TileAddr<P> ta = tile_addr<P>(io, tile, lane, s_ptr[w]);
    tile_in<P>(L, [&](int pp) { return ta.src(pp); }, lane);
    wave_sync();
    if (!INV) {
      pass_m32<kFormInter>(Pp, h, a1, bias1a, bias1b);
      wave_sync();
      pass_c32<FORM>(Pp, h, [&](int m) { return a2[m]; }, s_bias);
    } else {
      pass_c32<kFormInter>(Pp, h, [&](int m) { return a2[m]; }, s_bias);
      wave_sync();
      pass_m32<FORM>(Pp, h, a1, bias1a, bias1b);
    }
    wave_sync();
    tile_out<P>(L, [&](int pp) { return ta.dst(pp); }, lane);
    wave_sync();
  }
}

template <int SHAPE, int FORM, bool INV, class Tiles>
void launch_one(dim3 grid, int waves, const Tiles &io, const DeviceTables *tab, hipStream_t s)
{
  if (SHAPE == 16) {
    if (waves == 2)
      hipLaunchKernelGGL((k_mfma16<FORM, INV, 2, Tiles>), grid, dim3(128), 0, s, io, tab);
    else
      hipLaunchKernelGGL((k_mfma16<FORM, INV, 4, Tiles>), grid, dim3(256), 0, s, io, tab);
  } else {
    if (waves == 2)
      hipLaunchKernelGGL((k_mfma32<FORM, INV, 2, Tiles>), grid, dim3(128), 0, s, io, tab);
    else
      hipLaunchKernelGGL((k_mfma32<FORM, INV, 4, Tiles>), grid, dim3(256), 0, s, io, tab);
  }
}

template <int SHAPE, bool INV, class Tiles>
int launch_form(int form, dim3 grid, int waves, const Tiles &io, const DeviceTables *tab,
                hipStream_t s)
{
  switch (form) {
    case kFormCanon: launch_one<SHAPE, kFormCanon, INV>(grid, waves, io, tab, s); return 0;
    case kFormCentred: launch_one<SHAPE, kFormCentred, INV>(grid, waves, io, tab, s); return 0;
    default: return -1;
  }
}

template <class Tiles>
int launch_any(int shape, bool inverse, int form, const Tiles &io, const DeviceTables *tab,
               int grid_cap, int waves, hipStream_t stream)
{
  if (waves != 2 && waves != 4)
    return -1;
  const int P = shape == 32 ? 32 : 16;
  const size_t ntiles = (io.npoly + P - 1) / P;
  const size_t need = (ntiles + waves - 1) / waves;
  const dim3 grid((unsigned)(need < (size_t)grid_cap ? need : (size_t)grid_cap));
  if (shape == 32)
    return inverse ? launch_form<32, true>(form, grid, waves, io, tab, stream)
                   : launch_form<32, false>(form, grid, waves, io, tab, stream);
  return inverse ? launch_form<16, true>(form, grid, waves, io, tab, stream)
                 : launch_form<16, false>(form, grid, waves, io, tab, stream);
}

}  // namespace

int launch_mfma(int shape, bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                const DeviceTables *tab, int grid_cap, int waves, hipStream_t stream)
{
  const LinearTiles io{in, out, npoly};
  return launch_any(shape, inverse, form, io, tab, grid_cap, waves, stream);
}

int launch_mfma_ragged(int shape, bool inverse, int form, const RaggedTiles &io,
                       const DeviceTables *tab, int grid_cap, int waves, hipStream_t stream)
{
  return launch_any(shape, inverse, form, io, tab, grid_cap, waves, stream);
}

}  // namespace tk
