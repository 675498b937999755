// Fused module matrix-vector product t = invntt(A_hat o ntt(s)) for ML-KEM
// ranks k = 2..4, one kernel: the k secret polys of 16 instances are staged
// and transformed in LDS, each output row is accumulated from A_hat streamed
// out of HBM, and inverse-transformed before it is written.  HBM traffic per
// instance is (k*k + 2k) * 512 bytes; nothing intermediate leaves the CU.
//
// Arithmetic.  The products are accumulated in the plain domain:
//   r0 = sum_j a0*b0 + gamma*(a1*b1),  r1 = sum_j a0*b1 + a1*b0   (mod q)
// per degree-1 pair with gamma = zetas[64+i] * 2^-16 (and -gamma for the odd
// pair); ref basemul (ref/ntt.c:139-146) yields 2^-16 times these, and
// invntt's final 1441 multiplies by 2^16 * 128^-1, so
// invntt(sum basemul) == plain NTT^-1 of the plain sum: the last pass uses the
// inverse matrix without the Montgomery factor (m16_inv1p).  A_hat may hold
// any int16 (offset by 10q >= 32768 to make it non-negative); the forward NTT
// of s is canonical, so every product and sum fits in uint32 and is reduced
// once per output coefficient.
#include "tk_launch.h"
#include "tk_mfma_core.h"

namespace tk {
namespace {

__device__ __forceinline__ uint32_t red32(uint32_t x)
{
  const uint32_t t = __umulhi(x, kBarrettM);
  return x - t * (uint32_t)kQ;  // [0, 2q) for every x < 2^32
}

__device__ __forceinline__ long ld_a(const uint32_t (*m)[64][2], int limb, int lane)
{
  return *reinterpret_cast<const long *>(m[limb][lane]);
}

template <int K>
__global__ __launch_bounds__(64) void k_matvec(int16_t *__restrict__ t,
                                               const int16_t *__restrict__ A,
                                               const int16_t *__restrict__ s, size_t ninst,
                                               const DeviceTables *__restrict__ tab)
{
  constexpr int P = 16;  // instances per tile
  __shared__ __align__(16) int16_t s_hat[K][P * kStride];
  __shared__ __align__(16) int16_t s_acc[P * kStride];
  __shared__ __align__(16) int32_t s_bf2[16][16];
  __shared__ __align__(16) int32_t s_bi2[16][16];
  __shared__ uint16_t s_gamma[64];
  const int lane = threadIdx.x, g = lane >> 4, tt = lane & 15;

  for (int i = lane; i < 256; i += 64) {
    s_bf2[i >> 4][i & 15] = tab->bias_fwd2[i >> 4][i & 15];
    s_bi2[i >> 4][i & 15] = tab->bias_inv2[i >> 4][i & 15];
  }
  s_gamma[lane] = tab->gamma[lane];
  __syncthreads();

  const long f1lo = ld_a(tab->m16_fwd1[0], 0, lane), f1hi = ld_a(tab->m16_fwd1[0], 1, lane);
  const long i1lo = ld_a(tab->m16_inv1p[0], 0, lane), i1hi = ld_a(tab->m16_inv1p[0], 1, lane);
  const v4i fb1 = *reinterpret_cast<const v4i *>(&tab->bias_fwd1[0][4 * g]);
  const v4i ib1 = *reinterpret_cast<const v4i *>(&tab->bias_inv1p[0][4 * g]);
  auto f2lo = [&](int m) { return ld_a(tab->m16_fwd2[m], 0, lane); };
  auto f2hi = [&](int m) { return ld_a(tab->m16_fwd2[m], 1, lane); };
  auto i2lo = [&](int m) { return ld_a(tab->m16_inv2[m], 0, lane); };
  auto i2hi = [&](int m) { return ld_a(tab->m16_inv2[m], 1, lane); };

  const size_t ntiles = (ninst + P - 1) / P;
  for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const size_t i0 = tile * P;
    const int nvalid = (int)(ninst - i0 < (size_t)P ? ninst - i0 : P);

    // s_hat[j] = ntt(s_j) of the tile's instances, canonical.
#pragma unroll
    for (int j = 0; j < K; j++)
      tile_in<P>(s_hat[j],
                 [&](int pp) -> const int16_t * {
                   return pp < nvalid ? s + ((i0 + pp) * K + j) * kN : nullptr;
                 },
                 lane);
    wave_sync();
#pragma unroll
    for (int j = 0; j < K; j++)
      pass_m16<kFormInter>(s_hat[j] + tt * kStride, g, f1lo, f1hi, fb1);
    wave_sync();
#pragma unroll
    for (int j = 0; j < K; j++)
      pass_c16<kFormCanon>(s_hat[j] + tt * kStride, g, f2lo, f2hi, s_bf2);
    wave_sync();

    for (int i = 0; i < K; i++) {
      // acc = sum_j A_hat[i][j] o s_hat[j], 8 chunks of 8 coefficients per lane.
#pragma unroll 2
      for (int it = 0; it < P / 2; it++) {
        const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
        uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (pp < nvalid) {
#pragma unroll
          for (int j = 0; j < K; j++) {
            union {
              v4i v;
              int16_t h[8];
            } a, b;
            a.v = ld_stream(reinterpret_cast<const v4i *>(A + (((i0 + pp) * K + i) * K + j) * kN) + ch);
            b.v = *reinterpret_cast<const v4i *>(s_hat[j] + pp * kStride + ch * 8);
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const uint32_t gam = s_gamma[2 * ch + (u >> 1)];
              const uint32_t gm = (u & 1) ? (uint32_t)kQ - gam : gam;
              const uint32_t a0 = (uint32_t)((int)a.h[2 * u] + 10 * kQ);
              const uint32_t a1 = (uint32_t)((int)a.h[2 * u + 1] + 10 * kQ);
              const uint32_t b0 = (uint16_t)b.h[2 * u], b1 = (uint16_t)b.h[2 * u + 1];
              const uint32_t p11 = red32(__umul24(a1, b1));
              acc[2 * u] += __umul24(a0, b0) + __umul24(gm, p11);
              acc[2 * u + 1] += __umul24(a0, b1) + __umul24(a1, b0);
            }
          }
        }
        uint32_t r[8];
#pragma unroll
        for (int e = 0; e < 8; e++)
          r[e] = red32(acc[e]);
        const v4i out = {(int)pack2(r[0], r[1]), (int)pack2(r[2], r[3]), (int)pack2(r[4], r[5]),
                         (int)pack2(r[6], r[7])};
        *reinterpret_cast<v4i *>(s_acc + pp * kStride + ch * 8) = out;
      }
      wave_sync();
      pass_c16<kFormInter>(s_acc + tt * kStride, g, i2lo, i2hi, s_bi2);
      wave_sync();
      pass_m16<kFormCentred>(s_acc + tt * kStride, g, i1lo, i1hi, ib1);
      wave_sync();
      tile_out<P>(s_acc,
                  [&](int pp) -> int16_t * {
                    return pp < nvalid ? t + ((i0 + pp) * K + i) * kN : nullptr;
                  },
                  lane);
      wave_sync();
    }
  }
}

}  // namespace

int launch_matvec(int k, int16_t *t, const int16_t *A, const int16_t *s, size_t ninst,
                  const DeviceTables *tab, int ncu, hipStream_t stream)
{
  const size_t ntiles = (ninst + 15) / 16;
  auto go = [&](auto kernel) {
    const size_t cap = (size_t)ncu * resident_per_cu((const void *)kernel, 64);
    const dim3 grid((unsigned)(ntiles < cap ? ntiles : cap));
    hipLaunchKernelGGL(kernel, grid, dim3(64), 0, stream, t, A, s, ninst, tab);
    return 0;
  };
  switch (k) {
    case 2: return go(k_matvec<2>);
    case 3: return go(k_matvec<3>);
    case 4: return go(k_matvec<4>);
    default: return -1;
  }
}

}  // namespace tk
