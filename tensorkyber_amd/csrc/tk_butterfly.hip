// Plain method: scalar Cooley-Tukey (forward) and Gentleman-Sande (inverse)
// butterflies with Montgomery / Barrett reduction, operation for operation
// the reference C transform (ref/ntt.c:80-95, :106-126), so the TK_FORM_REF
// output is the reference's own int16 values.
//
// Mapping: 16 lanes per polynomial, 4 polynomials per wave, each wave owns
// its LDS slab.  A poly is viewed as x[m][c] = r[16m + c].  Layers 1-4
// (len 128..16) only mix the m index, so lane c first holds column c (16
// values, twiddles zetas[1..15] uniform across the wave); layers 5-7
// (len 8..2) stay inside a 16-coefficient block, so after one LDS transpose
// lane i holds block i with its 7 twiddles zetas[16+i], [32+2i+s],
// [64+4i+s].  HBM traffic is one coalesced 16-byte read and write per lane
// and chunk.
#include "tk_device.h"
#include "tk_launch.h"

namespace tk {
namespace {

constexpr int kStride = 264;  // int16 per poly slot in LDS (512 + 16 bytes)
constexpr int kPolysPerWave = 4;
constexpr int kWaves = 4;

__device__ __forceinline__ void ct(int16_t &a, int16_t &b, int16_t z, int16_t zq)
{
  const int16_t t = fqmul_zq(b, z, zq);
  b = (int16_t)(a - t);
  a = (int16_t)(a + t);
}

__device__ __forceinline__ void gs(int16_t &a, int16_t &b, int16_t z, int16_t zq)
{
  const int16_t t = a;
  a = barrett16((int16_t)(t + b));
  b = (int16_t)(b - t);
  b = fqmul_zq(b, z, zq);
}

template <int FORM>
__device__ __forceinline__ int16_t finish(int16_t v)
{
  if (FORM == kFormCanon)
    return canon16(v);
  if (FORM == kFormCentred)
    return centre16(v);
  return v;
}

__device__ __forceinline__ void stage_in(int16_t *L, const int16_t *__restrict__ in, size_t p0,
                                         size_t npoly, int lane)
{
#pragma unroll
  for (int it = 0; it < kPolysPerWave * 32 / 64; it++) {
    const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
    v4i v = {0, 0, 0, 0};
    if (p0 + pp < npoly)
      v = ld_stream(reinterpret_cast<const v4i *>(in + (p0 + pp) * kN) + ch);
    *reinterpret_cast<v4i *>(L + pp * kStride + ch * 8) = v;
  }
}

__device__ __forceinline__ void stage_out(const int16_t *L, int16_t *__restrict__ out, size_t p0,
                                          size_t npoly, int lane)
{
#pragma unroll
  for (int it = 0; it < kPolysPerWave * 32 / 64; it++) {
    const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
    const v4i v = *reinterpret_cast<const v4i *>(L + pp * kStride + ch * 8);
    if (p0 + pp < npoly)
      st_stream(reinterpret_cast<v4i *>(out + (p0 + pp) * kN) + ch, v);
  }
}

template <int FORM>
__global__ __launch_bounds__(256) void k_ntt_butterfly(int16_t *__restrict__ out,
                                                       const int16_t *__restrict__ in,
                                                       size_t npoly,
                                                       const DeviceTables *__restrict__ tab)
{
  __shared__ __align__(16) int16_t s_poly[kWaves][kPolysPerWave * kStride];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pl = lane >> 4, col = lane & 15;
  int16_t *L = s_poly[w];
  int16_t *P = L + pl * kStride;

  // Per-lane twiddles of layers 5-7 for block `col`.
  int16_t z5 = tab->zetas[16 + col], z5q = tab->zetas_qinv[16 + col];
  int16_t z6[2], z6q[2], z7[4], z7q[4];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    z6[s] = tab->zetas[32 + 2 * col + s];
    z6q[s] = tab->zetas_qinv[32 + 2 * col + s];
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    z7[s] = tab->zetas[64 + 4 * col + s];
    z7q[s] = tab->zetas_qinv[64 + 4 * col + s];
  }

  const size_t ngroups = (npoly + kPolysPerWave - 1) / kPolysPerWave;
  for (size_t grp = (size_t)blockIdx.x * kWaves + w; grp < ngroups;
       grp += (size_t)gridDim.x * kWaves) {
    const size_t p0 = grp * kPolysPerWave;
    stage_in(L, in, p0, npoly, lane);
    wave_sync();

    // Layers 1-4 on column `col`: x[m] = r[16m + col].
    int16_t x[16];
#pragma unroll
    for (int m = 0; m < 16; m++)
      x[m] = P[16 * m + col];
    int k = 1;
#pragma unroll
    for (int half = 8; half >= 1; half >>= 1) {      // len = 16*half
#pragma unroll
      for (int start = 0; start < 16; start += 2 * half) {
        const int16_t z = tab->zetas[k], zq = tab->zetas_qinv[k];
        k++;
#pragma unroll
        for (int j = start; j < start + half; j++)
          ct(x[j], x[j + half], z, zq);
      }
    }
#pragma unroll
    for (int m = 0; m < 16; m++)
      P[16 * m + col] = x[m];
    wave_sync();

    // Layers 5-7 on block `col`: x[j] = r[16*col + j].
    int16_t *B = P + 16 * col;
#pragma unroll
    for (int j = 0; j < 16; j++)
      x[j] = B[j];
#pragma unroll
    for (int j = 0; j < 8; j++)
      ct(x[j], x[j + 8], z5, z5q);
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        ct(x[8 * s + j], x[8 * s + j + 4], z6[s], z6q[s]);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        ct(x[4 * s + j], x[4 * s + j + 2], z7[s], z7q[s]);
#pragma unroll
    for (int j = 0; j < 16; j++)
      B[j] = finish<FORM>(x[j]);
    wave_sync();
    stage_out(L, out, p0, npoly, lane);
    wave_sync();
  }
}

template <int FORM>
__global__ __launch_bounds__(256) void k_invntt_butterfly(int16_t *__restrict__ out,
                                                          const int16_t *__restrict__ in,
                                                          size_t npoly,
                                                          const DeviceTables *__restrict__ tab)
{
  __shared__ __align__(16) int16_t s_poly[kWaves][kPolysPerWave * kStride];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pl = lane >> 4, col = lane & 15;
  int16_t *L = s_poly[w];
  int16_t *P = L + pl * kStride;

  // Per-lane twiddles of the inverse layers len 2, 4, 8 for block `col`
  // (ref/ntt.c:111-114 counts k down from 127).
  int16_t z7[4], z7q[4], z6[2], z6q[2];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    z7[s] = tab->zetas[127 - 4 * col - s];
    z7q[s] = tab->zetas_qinv[127 - 4 * col - s];
  }
#pragma unroll
  for (int s = 0; s < 2; s++) {
    z6[s] = tab->zetas[63 - 2 * col - s];
    z6q[s] = tab->zetas_qinv[63 - 2 * col - s];
  }
  const int16_t z5 = tab->zetas[31 - col], z5q = tab->zetas_qinv[31 - col];
  const int16_t f = (int16_t)kInvScale;
  const int16_t fq = (int16_t)(uint16_t)((uint32_t)(kInvScale * kQinv) & 0xffffu);

  const size_t ngroups = (npoly + kPolysPerWave - 1) / kPolysPerWave;
  for (size_t grp = (size_t)blockIdx.x * kWaves + w; grp < ngroups;
       grp += (size_t)gridDim.x * kWaves) {
    const size_t p0 = grp * kPolysPerWave;
    stage_in(L, in, p0, npoly, lane);
    wave_sync();

    int16_t x[16];
    int16_t *B = P + 16 * col;
#pragma unroll
    for (int j = 0; j < 16; j++)
      x[j] = B[j];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        gs(x[4 * s + j], x[4 * s + j + 2], z7[s], z7q[s]);
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        gs(x[8 * s + j], x[8 * s + j + 4], z6[s], z6q[s]);
#pragma unroll
    for (int j = 0; j < 8; j++)
      gs(x[j], x[j + 8], z5, z5q);
#pragma unroll
    for (int j = 0; j < 16; j++)
      B[j] = x[j];
    wave_sync();

#pragma unroll
    for (int m = 0; m < 16; m++)
      x[m] = P[16 * m + col];
    int k = 15;
#pragma unroll
    for (int half = 1; half <= 8; half <<= 1) {      // len = 16*half
#pragma unroll
      for (int start = 0; start < 16; start += 2 * half) {
        const int16_t z = tab->zetas[k], zq = tab->zetas_qinv[k];
        k--;
#pragma unroll
        for (int j = start; j < start + half; j++)
          gs(x[j], x[j + half], z, zq);
      }
    }
#pragma unroll
    for (int m = 0; m < 16; m++)
      P[16 * m + col] = finish<FORM>(fqmul_zq(x[m], f, fq));
    wave_sync();
    stage_out(L, out, p0, npoly, lane);
    wave_sync();
  }
}

}  // namespace

int launch_butterfly(bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                     const DeviceTables *tab, int ncu, hipStream_t stream)
{
  const size_t ngroups = (npoly + kPolysPerWave - 1) / kPolysPerWave;
  const size_t need = (ngroups + kWaves - 1) / kWaves;
  const int block = 64 * kWaves;
  // Four times the resident workgroups, each looping over its groups: on
  // MI355X 2^20 polys take 222-226 us forward and 306-312 us inverse, against
  // 231 and 335 us for one resident set (two: 228-232 and 315-317; eight:
  // the same as four) and 264 and 350 us for a grid that covers the batch.
  constexpr int kOversub = 4;
  auto go = [&](auto kernel) {
    const size_t cap = (size_t)ncu * resident_per_cu((const void *)kernel, block) * kOversub;
    const dim3 grid((unsigned)(need < cap ? need : cap));
    hipLaunchKernelGGL(kernel, grid, dim3(block), 0, stream, out, in, npoly, tab);
    return 0;
  };
  switch (form) {
    case kFormCanon: return inverse ? go(k_invntt_butterfly<kFormCanon>) : go(k_ntt_butterfly<kFormCanon>);
    case kFormCentred: return inverse ? go(k_invntt_butterfly<kFormCentred>) : go(k_ntt_butterfly<kFormCentred>);
    case kFormRef: return inverse ? go(k_invntt_butterfly<kFormRef>) : go(k_ntt_butterfly<kFormRef>);
    default: return -1;
  }
}

}  // namespace tk
