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
#include "tk_device.h"
#include "tk_launch.h"

namespace tk {
namespace {

constexpr int kStride = 264;  // int16 per poly slot in LDS (512 + 16 bytes)

// ---------------------------------------------------------------------------
// Tile staging: P polys, coalesced 16-byte global accesses, one LDS slab.
// `src` yields the address of tile poly pp (nullptr = beyond the batch).
// ---------------------------------------------------------------------------
template <int P, class Addr>
__device__ __forceinline__ void tile_in(int16_t *L, Addr src, int lane)
{
  v4i v[P / 2];
#pragma unroll
  for (int it = 0; it < P / 2; it++) {
    const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
    const int16_t *p = src(pp);
    v[it] = p ? reinterpret_cast<const v4i *>(p)[ch] : v4i{0, 0, 0, 0};
  }
#pragma unroll
  for (int it = 0; it < P / 2; it++) {
    const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
    *reinterpret_cast<v4i *>(L + pp * kStride + ch * 8) = v[it];
  }
}

template <int P, class Addr>
__device__ __forceinline__ void tile_out(const int16_t *L, Addr dst, int lane)
{
#pragma unroll
  for (int it = 0; it < P / 2; it++) {
    const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
    const v4i v = *reinterpret_cast<const v4i *>(L + pp * kStride + ch * 8);
    int16_t *p = dst(pp);
    if (p)
      reinterpret_cast<v4i *>(p)[ch] = v;
  }
}

template <int MODE>
__device__ __forceinline__ uint32_t epilogue(int lo, int hi)
{
  if (MODE == kFormInter)
    return reduce_2q(lo, hi);
  if (MODE == kFormCanon)
    return reduce_q(lo, hi);
  return reduce_q(lo, hi) - (uint32_t)kCentre;  // bias carried +kCentre
}

// ---------------------------------------------------------------------------
// 16x16x32 passes: lane l = 16g + t works on poly t of the tile; its K group
// and its output rows are the index set {4g, .., 4g+3}.
// ---------------------------------------------------------------------------

// Contraction over m for each column c (one matrix for all columns).
template <int MODE>
__device__ __forceinline__ void pass_m16(int16_t *P, int g, long alo, long ahi, v4i bias)
{
  const v4i zero = {0, 0, 0, 0};
#pragma unroll 4
  for (int c = 0; c < 16; c++) {
    const int16_t *col = P + 64 * g + c;
    const uint32_t d0 = (uint16_t)col[0] | ((uint32_t)(uint16_t)col[16] << 16);
    const uint32_t d1 = (uint16_t)col[32] | ((uint32_t)(uint16_t)col[48] << 16);
    const long b = digits8(d0, d1);
    const v4i lo = mfma16(alo, b, bias);
    const v4i hi = mfma16(ahi, b, zero);
    int16_t *o = P + 64 * g + c;
#pragma unroll
    for (int r = 0; r < 4; r++)
      o[16 * r] = (int16_t)epilogue<MODE>(lo[r], hi[r]);
  }
}

// Contraction over c inside each block m (one matrix per block).
template <int MODE>
__device__ __forceinline__ void pass_c16(int16_t *P, int g, const long *alo, const long *ahi,
                                         const int32_t (*s_bias)[16])
{
  const v4i zero = {0, 0, 0, 0};
#pragma unroll
  for (int m = 0; m < 16; m++) {
    uint2 *slot = reinterpret_cast<uint2 *>(P + 16 * m + 4 * g);
    const uint2 d = *slot;
    const long b = digits8(d.x, d.y);
    const v4i bias = *reinterpret_cast<const v4i *>(&s_bias[m][4 * g]);
    const v4i lo = mfma16(alo[m], b, bias);
    const v4i hi = mfma16(ahi[m], b, zero);
    uint32_t r[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
      r[i] = epilogue<MODE>(lo[i], hi[i]);
    *slot = make_uint2(pack2(r[0], r[1]), pack2(r[2], r[3]));
  }
}

// ---------------------------------------------------------------------------
// 32x32x32 passes: lane l = 32h + t works on poly t; its K group and output
// rows are U_h = {4h..4h+3, 8+4h..8+4h+3}; accumulator register q (< 8) is
// the lo limb of row U_h[q], register q+8 its hi limb.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int u32set(int h, int q) { return q < 4 ? 4 * h + q : 8 + 4 * h + q - 4; }

template <int MODE>
__device__ __forceinline__ void pass_m32(int16_t *P, int h, v4i a, const v4i &b0, const v4i &b1)
{
#pragma unroll 2
  for (int c = 0; c < 16; c++) {
    uint32_t d[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      d[j] = (uint16_t)P[16 * u32set(h, 2 * j) + c] |
             ((uint32_t)(uint16_t)P[16 * u32set(h, 2 * j + 1) + c] << 16);
    const v4i b = {(int)digit_lo(d[0], d[1]), (int)digit_lo(d[2], d[3]), (int)digit_hi(d[0], d[1]),
                   (int)digit_hi(d[2], d[3])};
    const v16i acc0 = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3],
                       0, 0, 0, 0, 0, 0, 0, 0};
    const v16i acc = mfma32(a, b, acc0);
#pragma unroll
    for (int q = 0; q < 8; q++)
      P[16 * u32set(h, q) + c] = (int16_t)epilogue<MODE>(acc[q], acc[q + 8]);
  }
}

template <int MODE>
__device__ __forceinline__ void pass_c32(int16_t *P, int h, const v4i *a,
                                         const int32_t (*s_bias)[16])
{
#pragma unroll
  for (int m = 0; m < 16; m++) {
    uint2 *s0 = reinterpret_cast<uint2 *>(P + 16 * m + 4 * h);
    uint2 *s1 = reinterpret_cast<uint2 *>(P + 16 * m + 8 + 4 * h);
    const uint2 x0 = *s0, x1 = *s1;
    const v4i b = {(int)digit_lo(x0.x, x0.y), (int)digit_lo(x1.x, x1.y), (int)digit_hi(x0.x, x0.y),
                   (int)digit_hi(x1.x, x1.y)};
    const v4i c0 = *reinterpret_cast<const v4i *>(&s_bias[m][4 * h]);
    const v4i c1 = *reinterpret_cast<const v4i *>(&s_bias[m][8 + 4 * h]);
    const v16i acc0 = {c0[0], c0[1], c0[2], c0[3], c1[0], c1[1], c1[2], c1[3],
                       0, 0, 0, 0, 0, 0, 0, 0};
    const v16i acc = mfma32(a[m], b, acc0);
    uint32_t r[8];
#pragma unroll
    for (int q = 0; q < 8; q++)
      r[q] = epilogue<MODE>(acc[q], acc[q + 8]);
    *s0 = make_uint2(pack2(r[0], r[1]), pack2(r[2], r[3]));
    *s1 = make_uint2(pack2(r[4], r[5]), pack2(r[6], r[7]));
  }
}

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

struct LinearTile {
  const int16_t *in;
  int16_t *out;
  size_t npoly;
};

template <int FORM, bool INV, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_mfma16(LinearTile io,
                                                        const DeviceTables *__restrict__ tab)
{
  constexpr int P = 16;
  __shared__ __align__(16) int16_t s_poly[WAVES][P * kStride];
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
    const size_t p0 = tile * P;
    tile_in<P>(L, [&](int pp) { return p0 + pp < io.npoly ? io.in + (p0 + pp) * kN : nullptr; },
               lane);
    wave_sync();
    if (!INV) {
      pass_m16<kFormInter>(Pp, g, a1lo, a1hi, bias1);
      wave_sync();
      pass_c16<FORM>(Pp, g, a2lo, a2hi, s_bias);
    } else {
      pass_c16<kFormInter>(Pp, g, a2lo, a2hi, s_bias);
      wave_sync();
      pass_m16<FORM>(Pp, g, a1lo, a1hi, bias1);
    }
    wave_sync();
    tile_out<P>(L, [&](int pp) { return p0 + pp < io.npoly ? io.out + (p0 + pp) * kN : nullptr; },
                lane);
    wave_sync();
  }
}

template <int FORM, bool INV, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_mfma32(LinearTile io,
                                                        const DeviceTables *__restrict__ tab)
{
  constexpr int P = 32;
  __shared__ __align__(16) int16_t s_poly[WAVES][P * kStride];
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
    const size_t p0 = tile * P;
    tile_in<P>(L, [&](int pp) { return p0 + pp < io.npoly ? io.in + (p0 + pp) * kN : nullptr; },
               lane);
    wave_sync();
    if (!INV) {
      pass_m32<kFormInter>(Pp, h, a1, bias1a, bias1b);
      wave_sync();
      pass_c32<FORM>(Pp, h, a2, s_bias);
    } else {
      pass_c32<kFormInter>(Pp, h, a2, s_bias);
      wave_sync();
      pass_m32<FORM>(Pp, h, a1, bias1a, bias1b);
    }
    wave_sync();
    tile_out<P>(L, [&](int pp) { return p0 + pp < io.npoly ? io.out + (p0 + pp) * kN : nullptr; },
                lane);
    wave_sync();
  }
}

template <int SHAPE, int FORM, bool INV>
void launch_one(dim3 grid, int waves, LinearTile io, const DeviceTables *tab, hipStream_t s)
{
  if (SHAPE == 16) {
    if (waves == 2)
      hipLaunchKernelGGL((k_mfma16<FORM, INV, 2>), grid, dim3(128), 0, s, io, tab);
    else
      hipLaunchKernelGGL((k_mfma16<FORM, INV, 4>), grid, dim3(256), 0, s, io, tab);
  } else {
    if (waves == 2)
      hipLaunchKernelGGL((k_mfma32<FORM, INV, 2>), grid, dim3(128), 0, s, io, tab);
    else
      hipLaunchKernelGGL((k_mfma32<FORM, INV, 4>), grid, dim3(256), 0, s, io, tab);
  }
}

template <int SHAPE, bool INV>
int launch_form(int form, dim3 grid, int waves, LinearTile io, const DeviceTables *tab,
                hipStream_t s)
{
  switch (form) {
    case kFormCanon: launch_one<SHAPE, kFormCanon, INV>(grid, waves, io, tab, s); return 0;
    case kFormCentred: launch_one<SHAPE, kFormCentred, INV>(grid, waves, io, tab, s); return 0;
    default: return -1;
  }
}

}  // namespace

int launch_mfma(int shape, bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                const DeviceTables *tab, int grid_cap, int waves, hipStream_t stream)
{
  const int P = shape == 32 ? 32 : 16;
  const size_t ntiles = (npoly + P - 1) / P;
  const size_t need = (ntiles + waves - 1) / waves;
  const dim3 grid((unsigned)(need < (size_t)grid_cap ? need : (size_t)grid_cap));
  const LinearTile io{in, out, npoly};
  if (shape == 32)
    return inverse ? launch_form<32, true>(form, grid, waves, io, tab, stream)
                   : launch_form<32, false>(form, grid, waves, io, tab, stream);
  return inverse ? launch_form<16, true>(form, grid, waves, io, tab, stream)
                 : launch_form<16, false>(form, grid, waves, io, tab, stream);
}

}  // namespace tk
