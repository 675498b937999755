// Fused module matrix-vector product t = invntt(A_hat o ntt(s)) for ML-KEM
// ranks k = 2..4, one kernel.  A workgroup has k waves and works on tiles of
// 16 instances:
//
//   1. wave j stages s_j of the tile's 16 instances into its LDS slab,
//      forward-transforms it (canonical) and tabulates gamma * b1 for its
//      pairs;                                                   -- barrier --
//   2. wave i accumulates output row i, sum_j A_hat[i][j] o s_hat[j], with
//      A_hat streamed from HBM and every s_hat[j] read from LDS; the row is
//      held in registers;                                       -- barrier --
//   3. wave i writes its row into its own slab, inverse-transforms it and
//      stores t_i.
//
// HBM traffic per instance is (k*k + 2k) * 512 bytes; nothing intermediate
// leaves the CU.
//
// Arithmetic.  The products are accumulated in the plain domain:
//   r0 = sum_j a0*b0 + a1*(gamma*b1),  r1 = sum_j a0*b1 + a1*b0   (mod q)
// per degree-1 pair with gamma = zetas[64+i] * 2^-16 (and -gamma for the odd
// pair of group i).  Ref basemul (ref/ntt.c:139-146) yields 2^-16 times these
// and invntt's final 1441 multiplies by 2^16 * 128^-1, so
// invntt(sum basemul) == plain NTT^-1 of the plain sum: the last pass uses
// the inverse matrix without the Montgomery factor (m16_inv1p).
//
// Each sum is two v_dot2_i32_i16 per pair and source j: A_hat's packed pair
// (a0, a1) against (b0, gamma*b1 mod 2q) and against (b1, b0).  A_hat may hold
// any int16 and b, gamma*b1 lie in [0, 2q), so |sum| <= k * 32768 * (q-1 +
// 2q-1) < 1.31e9 < 2^31 for k <= 4; it is reduced once per output
// coefficient.
#include <type_traits>

#include "tk_launch.h"
#include "tk_mfma_core.h"

namespace tk {
namespace {

constexpr int kTileInst = 16;  // instances per tile: the MFMA16 poly width
constexpr int kBufferRsrcWord3 = 0x00020000;  // gfx9 buffer descriptor: 32-bit data format
// Multiple of q above every |sum|: 4 * 32768 * (3328 + 6657) < kSumOffset.
constexpr uint32_t kSumOffset = (4u * 32768u * (3328u + 6657u) / 3329u + 1u) * 3329u;

// [0, 2q) representative of x < 2^24.
__device__ __forceinline__ uint32_t red24(uint32_t x)
{
  return mad_i24(mulhi_u24(x, kBarrettM), -kQ, x);
}

// [0, 2q) representative of x < 2^32 (quotient estimate low by at most 1).
__device__ __forceinline__ uint32_t red32(uint32_t x)
{
  return mad_i24(__umulhi(x, kBarrettM), -kQ, x);
}

__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int c)
{
  typedef short v2s __attribute__((ext_vector_type(2)));
  return __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, a), __builtin_bit_cast(v2s, b), c, false);
}

__device__ __forceinline__ long ld_a(const uint32_t (*m)[64][2], int limb, int lane)
{
  return *reinterpret_cast<const long *>(m[limb][lane]);
}

// SBITS = 0: s holds int16 polys [ninst][K][256]; SBITS = 2, 3, 4: s holds
// the packed small-coefficient form [ninst][K][packed bytes].
template <int K, int SBITS>
__global__ __launch_bounds__(64 * K) void k_matvec(int16_t *__restrict__ t,
                                                   const int16_t *__restrict__ A,
                                                   const void *__restrict__ s, size_t ninst,
                                                   const DeviceTables *__restrict__ tab)
{
  constexpr int P = kTileInst;
  __shared__ __align__(16) int16_t s_slab[K][P * kStride];
  __shared__ __align__(16) uint32_t s_g[K][P][64];  // (gamma*b1, (q-gamma)*b3) per group
  __shared__ __align__(16) int32_t s_bf2[16][16];
  __shared__ __align__(16) int32_t s_bi2[16][16];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, tt = lane & 15;

  for (int i = threadIdx.x; i < 256; i += 64 * K) {
    s_bf2[i >> 4][i & 15] = tab->bias_fwd2[i >> 4][i & 15];
    s_bi2[i >> 4][i & 15] = tab->bias_inv2[i >> 4][i & 15];
  }
  __syncthreads();

  const long f1lo = ld_a(tab->m16_fwd1[0], 0, lane), f1hi = ld_a(tab->m16_fwd1[0], 1, lane);
  const long i1lo = ld_a(tab->m16_inv1p[0], 0, lane), i1hi = ld_a(tab->m16_inv1p[0], 1, lane);
  const v4i fb1 = *reinterpret_cast<const v4i *>(&tab->bias_fwd1[0][4 * g]);
  const v4i ib1 = *reinterpret_cast<const v4i *>(&tab->bias_inv1p[0][4 * g]);
  // The per-block matrices of both directions are fetched at each use from
  // the (cache-resident) table through a buffer descriptor: one lane-offset
  // register plus scalar offsets, where held addresses or values would cost
  // 128 registers.
  const __amdgpu_buffer_rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<DeviceTables *>(tab), 0, (int)sizeof(DeviceTables), kBufferRsrcWord3);
  const int loff = lane * 8;
  auto mat = [&](int off) {
    return __builtin_bit_cast(long, __builtin_amdgcn_raw_buffer_load_b64(trs, loff, off, 0));
  };
  constexpr int kF2 = (int)offsetof(DeviceTables, m16_fwd2), kI2 = (int)offsetof(DeviceTables, m16_inv2);
  auto f2lo = [&](int m) { return mat(kF2 + m * 1024); };
  auto f2hi = [&](int m) { return mat(kF2 + m * 1024 + 512); };
  auto i2lo = [&](int m) { return mat(kI2 + m * 1024); };
  auto i2hi = [&](int m) { return mat(kI2 + m * 1024 + 512); };
  // Lane `lane` tabulates group i = lane: pairs (4i, 4i+1) and (4i+2, 4i+3).
  const uint32_t gam = tab->gamma[lane], gam_neg = (uint32_t)kQ - gam;

  int16_t *const slab = s_slab[w];
  const size_t ntiles = (ninst + P - 1) / P;
  for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const size_t i0 = tile * P;
    const int nvalid = (int)(ninst - i0 < (size_t)P ? ninst - i0 : P);

    // 1. s_hat[w] = ntt(s_w), canonical, and its gamma products.
    if constexpr (SBITS == 0) {
      tile_in<P>(slab,
                 [&](int pp) -> const int16_t * {
                   return pp < nvalid ? static_cast<const int16_t *>(s) + ((i0 + pp) * K + w) * kN
                                      : nullptr;
                 },
                 lane);
    } else {
      constexpr size_t PPB = packed_poly_bytes<SBITS>();
      tile_in_packed<P, SBITS>(slab,
                               [&](int pp) -> const uint8_t * {
                                 return pp < nvalid ? static_cast<const uint8_t *>(s) +
                                                          ((i0 + pp) * K + w) * PPB
                                                    : nullptr;
                               },
                               lane);
    }
    wave_sync();
    pass_m16<kFormInter>(slab + tt * kStride, g, f1lo, f1hi, fb1);
    wave_sync();
    pass_c16<kFormCanon>(slab + tt * kStride, g, f2lo, f2hi, s_bf2);
    wave_sync();
#pragma unroll 4
    for (int pp = 0; pp < P; pp++) {
      const uint2 d = ld2(slab + pp * kStride + 4 * lane);
      s_g[w][pp][lane] = pack2(red24((d.x >> 16) * gam), red24((d.y >> 16) * gam_neg));
    }
    __syncthreads();

    // 2. Row w: lane covers instance pp, coefficients 8ch .. 8ch+7.  The
    // A_hat loads of the next chunk are in flight while this one is summed;
    // the scheduling barrier keeps the compiler from hoisting more of them.
    auto load_row = [&](v4i *a, int it) {
      const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
      const v4i *arow =
          reinterpret_cast<const v4i *>(A + ((i0 + pp) * K + w) * (size_t)(K * kN)) + ch;
#pragma unroll
      for (int j = 0; j < K; j++)
        a[j] = pp < nvalid ? ld_stream(arow + j * (kN / 8)) : v4i{0, 0, 0, 0};
    };
    uint32_t res[P / 2][4];
    v4i a_cur[K], a_nxt[K];
    load_row(a_cur, 0);
#pragma unroll
    for (int it = 0; it < P / 2; it++) {
      const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
      if (it + 1 < P / 2)
        load_row(a_nxt, it + 1);
      int r0[4] = {0, 0, 0, 0}, r1[4] = {0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < K; j++) {
        const v4i b = *reinterpret_cast<const v4i *>(s_slab[j] + pp * kStride + ch * 8);
        const uint2 gg = *reinterpret_cast<const uint2 *>(&s_g[j][pp][2 * ch]);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t bu = (uint32_t)b[u], gs = u < 2 ? gg.x : gg.y;
          const uint32_t x = __builtin_amdgcn_perm(gs, bu, (u & 1) ? 0x07060100u : 0x05040100u);
          const uint32_t y = __builtin_amdgcn_alignbit(bu, bu, 16);
          r0[u] = dot2((uint32_t)a_cur[j][u], x, r0[u]);
          r1[u] = dot2((uint32_t)a_cur[j][u], y, r1[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        res[it][u] = pack2(red32((uint32_t)r0[u] + kSumOffset), red32((uint32_t)r1[u] + kSumOffset));
#pragma unroll
      for (int j = 0; j < K; j++)
        a_cur[j] = a_nxt[j];
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();  // every wave is done reading the slabs

    // 3. t_w = invntt(row w), centred.
#pragma unroll
    for (int it = 0; it < P / 2; it++) {
      const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
      *reinterpret_cast<v4i *>(slab + pp * kStride + ch * 8) =
          v4i{(int)res[it][0], (int)res[it][1], (int)res[it][2], (int)res[it][3]};
    }
    wave_sync();
    pass_c16<kFormInter>(slab + tt * kStride, g, i2lo, i2hi, s_bi2);
    wave_sync();
    pass_m16<kFormCentred>(slab + tt * kStride, g, i1lo, i1hi, ib1);
    wave_sync();
    tile_out<P>(slab,
                [&](int pp) -> int16_t * {
                  return pp < nvalid ? t + ((i0 + pp) * K + w) * kN : nullptr;
                },
                lane);
    wave_sync();
  }
}

}  // namespace

int launch_matvec(int k, int s_bits, int16_t *t, const int16_t *A, const void *s, size_t ninst,
                  const DeviceTables *tab, int ncu, hipStream_t stream)
{
  const size_t ntiles = (ninst + kTileInst - 1) / kTileInst;
  auto go = [&](auto kernel, int waves) {
    const size_t cap = (size_t)ncu * resident_per_cu((const void *)kernel, 64 * waves);
    const dim3 grid((unsigned)(ntiles < cap ? ntiles : cap));
    hipLaunchKernelGGL(kernel, grid, dim3(64 * waves), 0, stream, t, A, s, ninst, tab);
    return 0;
  };
  auto with_k = [&](auto kbits) {
    constexpr int B = decltype(kbits)::value;
    switch (k) {
      case 2: return go(k_matvec<2, B>, 2);
      case 3: return go(k_matvec<3, B>, 3);
      case 4: return go(k_matvec<4, B>, 4);
      default: return -1;
    }
  };
  switch (s_bits) {
    case 0: return with_k(std::integral_constant<int, 0>{});
    case 2: return with_k(std::integral_constant<int, 2>{});
    case 3: return with_k(std::integral_constant<int, 3>{});
    case 4: return with_k(std::integral_constant<int, 4>{});
    default: return -1;
  }
}

}  // namespace tk
