// Fused module matrix-vector product t = invntt(A_hat o ntt(s)) for ML-KEM
// ranks k = 2..4, one kernel.  A workgroup has k waves and works on tiles of
// 16 instances:
//
//   1. wave j stages s_j of the tile's 16 instances into its LDS slab and
//      forward-transforms it (canonical);                        -- barrier --
//   2. wave i accumulates output row i, sum_j A_hat[i][j] o s_hat[j], with
//      A_hat streamed from HBM and every s_hat[j] read from LDS (gamma * b1
//      formed as it is read); the row is held in registers;     -- barrier --
//   3. wave i writes its row into its own slab, inverse-transforms it and
//      stores t_i.
//
// HBM traffic per instance is (k*k + 2k) * 512 bytes; nothing intermediate
// leaves the CU.
//
// The same kernel, by MODE, runs the K-PKE arithmetic cores of ML-KEM as the
// pqcrystals-kyber reference composes them (ref/indcpa.c, upstream of
// ref/), each output ending in that reference's poly_reduce (barrett_reduce,
// ref/reduce.c:35-42), so outputs are its raw int16 values:
//   keygen:  s_hat = poly_ntt(s) is written in step 1; step 3 stages e_i,
//            forward-transforms it and writes t_hat_i = reduce(row_i +
//            e_hat_i) in the NTT domain (no inverse);
//   encrypt: k+1 waves and rows: u_i from row i of the transposed matrix,
//            v from t_hat; step 3 inverse-transforms and adds e1_i, or e2 and
//            m;
//   decrypt: w = reduce(v - invntt(s_hat . ntt(u))) has one row, so it
//            runs in its own kernel (k_decrypt below): one wave per tile
//            forms the row alone and needs no workgroup barrier.
// The adds are exact mod q.  The reference's poly_add / poly_sub are int16
// adds on invntt's (-q, q) representatives; they agree with these wherever
// they do not overflow int16, which holds for every ML-KEM input (noise of
// at most eta = 3, message coefficients 0 or (q+1)/2, decompressed u and v
// in [0, q)).
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
// (a0, a1) against (b0, gamma*b1 mod 2q) and against (b1, b0).  The forward
// transforms leave b in [0, 2q) (no final fold: gamma * b1 < 2q * q is then
// reduced with the 32-bit Barrett step), A_hat may hold any int16 and
// gamma*b1 lies in [0, 2q), so |sum| <= k * 32768 * (2q-1 + 2q-1) < 1.75e9 <
// 2^31 for k <= 4; it is reduced once per output coefficient.
#include <type_traits>

#include "tk_launch.h"
#include "tk_mfma_core.h"

namespace tk {
namespace {

constexpr int kTileInst = 16;  // instances per tile: the MFMA16 poly width
// Multiple of q above every |sum|: 4 * 32768 * (6657 + 6657) < kSumOffset,
// and kSumOffset + |sum| < 2^32 (red32's domain).
constexpr uint32_t kSumOffset = (4u * 32768u * (6657u + 6657u) / 3329u + 1u) * 3329u;
static_assert((uint64_t)kSumOffset + 4ull * 32768ull * 13314ull < (1ull << 32), "red32 domain");

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

// barrett_reduce of each int16 lane of a v4i (centred representatives).
__device__ __forceinline__ v4i barrett_v4(const v4i x)
{
  v4i r;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int16_t lo = (int16_t)(x[d] & 0xffff), hi = (int16_t)(x[d] >> 16);
    r[d] = (int)pack2((uint16_t)barrett16(lo), (uint16_t)barrett16(hi));
  }
  return r;
}

// Lane-wise int16 add (no wrap for the operands it is used on: < 8q).  The
// elements are copied to scalars first: __builtin_bit_cast of a vector
// element lvalue reads element 0 here, whatever the index.
__device__ __forceinline__ v4i add16(const v4i a, const v4i b)
{
  typedef short v2s __attribute__((ext_vector_type(2)));
  v4i r;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int ad = a[d], bd = b[d];
    r[d] = __builtin_bit_cast(int, __builtin_bit_cast(v2s, ad) + __builtin_bit_cast(v2s, bd));
  }
  return r;
}

// Centred representative of SX*x + y + z per int16 lane, exact mod q for any
// int16 operands (|sum| < 21q is lifted into [0, 2^24) and reduced).
template <int SX>
__device__ __forceinline__ v4i centred_sum(const v4i x, const v4i y, const v4i z)
{
  auto one = [](int a) -> uint32_t {
    uint32_t r = red24((uint32_t)(a + 21 * kQ));  // [0, 2q)
    r = min(r, r - (uint32_t)kQ);                  // [0, q)
    return r > (uint32_t)kCentre ? r - (uint32_t)kQ : r;
  };
  v4i out;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int xl = (int16_t)(x[d] & 0xffff), xh = x[d] >> 16;
    const int yl = (int16_t)(y[d] & 0xffff), yh = y[d] >> 16;
    const int zl = (int16_t)(z[d] & 0xffff), zh = z[d] >> 16;
    out[d] = (int)pack2(one(SX * xl + yl + zl), one(SX * xh + yh + zh));
  }
  return out;
}

template <int MODE, int K>
constexpr int fused_waves()
{
  return MODE == kFusedEncrypt ? K + 1 : K;
}

template <int MODE, int K>
constexpr int fused_rows()
{
  return MODE == kFusedEncrypt ? K + 1 : K;
}

// Buffer roles of FusedIo / FusedInst, in FusedInst's order.
enum Role { rOut0 = 0, rOut1, rMat, rMat2, rX, rAdd0, rAdd1, rAdd2 };

// Per-instance stride of each role in the contiguous layout, in int16.
template <int MODE, int K>
__device__ __forceinline__ size_t role_stride(int role)
{
  switch (role) {
    case rOut0: return MODE == kFusedDecrypt ? kN : K * kN;
    case rOut1: return MODE == kFusedEncrypt ? kN : K * kN;
    case rMat: return MODE == kFusedDecrypt ? K * kN : K * K * kN;
    case rMat2:
    case rX: return K * kN;
    case rAdd0: return MODE == kFusedDecrypt ? kN : K * kN;
    default: return kN;
  }
}

__device__ __forceinline__ const int16_t *role_base(const FusedIo &io, int role)
{
  return role == rOut0   ? io.out0
         : role == rOut1 ? io.out1
         : role == rMat  ? io.mat
         : role == rMat2 ? io.mat2
         : role == rX    ? static_cast<const int16_t *>(io.x)
         : role == rAdd0 ? io.add0
         : role == rAdd1 ? io.add1
                         : io.add2;
}

// The polys of one buffer role for the 16 instances of a tile, in the
// staging lane order: chunk `it` of a tile covers instances 2*it and
// 2*it + 1 (lanes 0-31 and 32-63), 16 bytes per lane.  Poly `sub` of an
// instance is at + sub * 256 int16 from the instance's start.
//
// Contiguous layout (RoleTile<false>): instance i of the role starts at
// base + i * stride.  A buffer descriptor spans the tile's valid instances,
// so accesses of instances past the batch read zero or are dropped, with no
// per-lane branch; a role with no buffer gets an empty descriptor (reads
// zero).
// Table layout (RoleTile<true>): the instances' pointers parked in LDS;
// null pointers (instances past the batch, absent buffers) read zero and
// are not written.
template <bool TABLE>
struct RoleTile;

template <>
struct RoleTile<false> {
  __amdgpu_buffer_rsrc_t r;
  int voff;  // lane's byte offset within chunk 0
  int sb;    // per-instance stride in bytes
  __device__ __forceinline__ RoleTile(const int16_t *base, size_t i0, int nvalid, int stride,
                                      int lane)
      : r(span_rsrc(base ? base + i0 * (size_t)stride : nullptr,
                    base ? (uint32_t)nvalid * stride * 2 : 0u)),
        voff((lane >> 5) * stride * 2 + (lane & 31) * 16),
        sb(stride * 2)
  {
  }
  __device__ __forceinline__ v4i load(int it, int sub) const
  {
    return __builtin_bit_cast(
        v4i, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 2 * it * sb + sub * 512, kBufferNT));
  }
  // The whole offset in the VGPR operand (see span_rsrc's note on stores).
  __device__ __forceinline__ void store(int it, int sub, v4i v) const
  {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, v), r,
                                           voff + 2 * it * sb + sub * 512, 0, kBufferNT);
  }
};

template <>
struct RoleTile<true> {
  const int16_t *const (*ptrs)[8];  // [instance][role], in LDS
  int role, lane;
  __device__ __forceinline__ const int16_t *at(int it, int sub) const
  {
    const int16_t *p = ptrs[2 * it + (lane >> 5)][role];
    return p ? p + sub * kN + (lane & 31) * 8 : nullptr;
  }
  __device__ __forceinline__ v4i load(int it, int sub) const
  {
    const int16_t *p = at(it, sub);
    return p ? ld_stream(reinterpret_cast<const v4i *>(p)) : v4i{0, 0, 0, 0};
  }
  __device__ __forceinline__ void store(int it, int sub, v4i v) const
  {
    int16_t *p = const_cast<int16_t *>(at(it, sub));
    if (p)
      st_stream(reinterpret_cast<v4i *>(p), v);
  }
};

// Poly `sub` of the tile's instances into / out of the slab.  All loads (or
// LDS reads) are issued before the first use.
template <int P, class RT>
__device__ __forceinline__ void role_in(int16_t *slab, const RT &rt, int sub, int lane)
{
  v4i v[P / 2];
#pragma unroll
  for (int it = 0; it < P / 2; it++)
    v[it] = rt.load(it, sub);
#pragma unroll
  for (int it = 0; it < P / 2; it++) {
    const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
    *reinterpret_cast<v4i *>(slab + pp * kStride + ch * 8) = v[it];
  }
}

template <int P, class RT>
__device__ __forceinline__ void role_out(const int16_t *slab, const RT &rt, int sub, int lane)
{
  v4i v[P / 2];
#pragma unroll
  for (int it = 0; it < P / 2; it++) {
    const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
    v[it] = *reinterpret_cast<const v4i *>(slab + pp * kStride + ch * 8);
  }
#pragma unroll
  for (int it = 0; it < P / 2; it++)
    rt.store(it, sub, v[it]);
}

// K: module rank; MODE: kFusedMatvec / Keygen / Encrypt (FusedIo
// documents the buffers of each); the transformed vector holds int16 polys
// [ninst][K][256].  Contiguous matvec batches, int16 or packed, run
// k_matvec_wide instead; this kernel takes them from plans.  TABLE: each
// instance's buffers come from the per-instance table io.inst; the tile's
// 16 rows of pointers are parked in LDS, double-buffered by tile parity.
template <int K, int MODE, bool TABLE>
__global__ __launch_bounds__((64 * fused_waves<MODE, K>())) void k_fused(
    const FusedIo io, const DeviceTables *__restrict__ tab)
{
  constexpr int P = kTileInst, W = fused_waves<MODE, K>(), R = fused_rows<MODE, K>();
  __shared__ const int16_t *s_ip[TABLE ? 2 : 1][TABLE ? P : 1][8];
  __shared__ __align__(16) int16_t s_slab[W][P * kStride];
  __shared__ __align__(16) int32_t s_bf2[16][16];
  __shared__ __align__(16) int32_t s_bi2[16][16];
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, g = lane >> 4, tt = lane & 15;

  for (int i = threadIdx.x; i < 256; i += 64 * W) {
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
  // Lane (., ch) reads coefficients 8ch .. 8ch+7 of a poly in the row
  // products: pairs u = 0..3, pair u in group 2ch + u/2, with gamma for even
  // u and q - gamma for odd u.  gamma * b1 is formed where it is used: a
  // table of it per transformed poly would cost 4 KB of LDS per wave, and
  // with it a workgroup per CU.
  uint32_t gin[4];
  {
    const uint32_t ga = tab->gamma[2 * (lane & 31)], gb = tab->gamma[2 * (lane & 31) + 1];
    gin[0] = ga;
    gin[1] = (uint32_t)kQ - ga;
    gin[2] = gb;
    gin[3] = (uint32_t)kQ - gb;
  }

  int16_t *const slab = s_slab[w];
  int16_t *const pslab = slab + tt * kStride;
  // Accessor of a role for the current tile (i0, nvalid and the parity of
  // the tile are the loop's).
  size_t i0 = 0;
  int par = 0, nvalid = 0;
  auto role = [&](int rl) {
    if constexpr (TABLE)
      return RoleTile<true>{s_ip[par], rl, lane};
    else
      return RoleTile<false>(role_base(io, rl), i0, nvalid, (int)role_stride<MODE, K>(rl), lane);
  };
  // Matrix row of output row w: the role and the first of its K polys.
  const int row_role = (MODE == kFusedEncrypt && w == K) ? rMat2 : rMat;
  const int row_sub = (MODE == kFusedEncrypt && w == K) ? 0 : w * K;
  // Forward transform of the staged slab to canonical values.
  auto forward = [&]() {
    wave_sync();
    pass_m16<kFormInter>(pslab, g, f1lo, f1hi, fb1);
    wave_sync();
    pass_c16<kFormInter>(pslab, g, f2lo, f2hi, s_bf2);
    wave_sync();
  };

  const size_t ntiles = (io.ninst + P - 1) / P;
  for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x, par ^= (TABLE ? 1 : 0)) {
    i0 = tile * P;
    nvalid = (int)(io.ninst - i0 < (size_t)P ? io.ninst - i0 : P);
    if constexpr (TABLE) {
      // Buffer par was last read in the tile before the previous one, whose
      // readers have all passed the previous tile's barriers.
      if (w == 0)
        for (int idx = lane; idx < P * 8; idx += 64) {
          const int pp = idx >> 3, role = idx & 7;
          s_ip[par][pp][role] =
              pp < nvalid ? reinterpret_cast<const int16_t *const *>(io.inst + i0 + pp)[role]
                          : nullptr;
        }
      __syncthreads();
    }

    // 1. x_hat[w] = ntt(x_w), canonical, and its gamma products.
    if (w < K) {
      role_in<P>(slab, role(rX), w, lane);
      forward();
    }
    __syncthreads();

    // 2. Row w: lane covers instance pp, coefficients 8ch .. 8ch+7.  The
    // matrix loads of the next chunk are in flight while this one is summed;
    // the scheduling barrier keeps the compiler from hoisting more of them.
    uint32_t res[P / 2][4];
    if (w < R) {
      const auto mr = role(row_role);
      auto load_row = [&](v4i *a, int it) {
#pragma unroll
        for (int j = 0; j < K; j++)
          a[j] = mr.load(it, row_sub + j);
      };
      v4i a_cur[K], a_nxt[K];
      load_row(a_cur, 0);
#pragma unroll
      for (int it = 0; it < P / 2; it++) {
        const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
        if (it + 1 < P / 2)
          load_row(a_nxt, it + 1);
        if constexpr (MODE == kFusedKeygen) {
          if (it == P / 2 - 2) {
            // s_hat = poly_ntt(s_w), centred, from the wave's own slab (the
            // other waves only read it until the barrier), once the row's
            // last loads are out: vector memory operations complete in
            // order, so stores issued before a row chunk's loads would be
            // waited for with them.  65,536 instances: 124 -> 123.6 us at
            // k = 3, 188 -> 182 us at k = 4.
            const auto o1 = role(rOut1);
#pragma unroll
            for (int i2 = 0; i2 < P / 2; i2++) {
              const int q2 = i2 * 64 + lane, pp2 = q2 >> 5, ch2 = q2 & 31;
              o1.store(i2, w,
                       barrett_v4(*reinterpret_cast<const v4i *>(slab + pp2 * kStride + ch2 * 8)));
            }
          }
        }
        int r0[4] = {0, 0, 0, 0}, r1[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < K; j++) {
          const v4i b = *reinterpret_cast<const v4i *>(s_slab[j] + pp * kStride + ch * 8);
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const uint32_t bu = (uint32_t)b[u];
            const uint32_t x = __builtin_amdgcn_perm(red32((bu >> 16) * gin[u]), bu, 0x05040100u);
            const uint32_t y = __builtin_amdgcn_alignbit(bu, bu, 16);
            r0[u] = dot2((uint32_t)a_cur[j][u], x, r0[u]);
            r1[u] = dot2((uint32_t)a_cur[j][u], y, r1[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
          res[it][u] =
              pack2(red32((uint32_t)r0[u] + kSumOffset), red32((uint32_t)r1[u] + kSumOffset));
#pragma unroll
        for (int j = 0; j < K; j++)
          a_cur[j] = a_nxt[j];
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();  // every wave is done reading the slabs

    // 3. Output row w.
    if (w < R) {
      if constexpr (MODE == kFusedKeygen) {
        // t_hat_w = reduce(row + poly_ntt(e_w)), NTT domain.
        role_in<P>(slab, role(rAdd0), w, lane);
        forward();
        const auto o0 = role(rOut0);
#pragma unroll
        for (int it = 0; it < P / 2; it++) {
          const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
          const v4i e = *reinterpret_cast<const v4i *>(slab + pp * kStride + ch * 8);
          const v4i sum = add16(v4i{(int)res[it][0], (int)res[it][1], (int)res[it][2],
                                    (int)res[it][3]},
                                e);  // < 3q: no wrap
          o0.store(it, w, barrett_v4(sum));
        }
      } else {
#pragma unroll
        for (int it = 0; it < P / 2; it++) {
          const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
          *reinterpret_cast<v4i *>(slab + pp * kStride + ch * 8) =
              v4i{(int)res[it][0], (int)res[it][1], (int)res[it][2], (int)res[it][3]};
        }
        wave_sync();
        pass_c16<kFormInter>(pslab, g, i2lo, i2hi, s_bi2);
        wave_sync();
        pass_m16<kFormCentred>(pslab, g, i1lo, i1hi, ib1);
        wave_sync();
        if constexpr (MODE == kFusedMatvec) {
          role_out<P>(slab, role(rOut0), w, lane);
        } else {
          // Addends of the output row: e1_w (a u row), e2 and m (the v row),
          // four chunks at a time: each chunk's loads go out before its store
          // (outputs alias their addend exactly, chunk for chunk), and the
          // 32 registers of a group let encryption run 4 waves per SIMD
          // (all eight chunks at once took 147 VGPRs, 3 waves, and 1% longer).
          const bool vrow = MODE == kFusedEncrypt && w == K;
          const int sub = (MODE == kFusedEncrypt && !vrow) ? w : 0;
          const auto ya = role(vrow ? rAdd1 : rAdd0);
          const auto za = role(rAdd2);
          const auto oa = role(vrow ? rOut1 : rOut0);
          const v4i zero = {0, 0, 0, 0};
          constexpr int G = 4;
#pragma unroll
          for (int g0 = 0; g0 < P / 2; g0 += G) {
            v4i y[G], z[G];
#pragma unroll
            for (int i = 0; i < G; i++) {
              y[i] = ya.load(g0 + i, sub);
              z[i] = vrow ? za.load(g0 + i, 0) : zero;
            }
#pragma unroll
            for (int i = 0; i < G; i++) {
              const int q = (g0 + i) * 64 + lane, pp = q >> 5, ch = q & 31;
              const v4i x = *reinterpret_cast<const v4i *>(slab + pp * kStride + ch * 8);
              // u_w = reduce(x + e1_w), v = reduce(x + e2 + m)
              oa.store(g0 + i, sub, centred_sum<1>(x, y[i], z[i]));
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      wave_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// Wide matvec: the fused product (contiguous buffers, s int16 or packed) with one
// workgroup of several K-wave tiles per CU, so that the transform matrices
// of both directions fit once in LDS for all of them.  The transforms then
// issue no vector memory load, and since vector memory loads complete in
// order, loads issued before a transform stay in flight through it instead
// of being waited for at its first matrix fetch:
//   - the first kWidePre chunks of each wave's matrix row go out before its
//     forward transform (phase 1) and are ready when phase 2 starts; phase 2
//     keeps kWidePre chunks in flight;
//   - the next tile's s_w goes out before the inverse transform (phase 3),
//     into the registers the row results free.
// The tiles of a workgroup go through the phases together (workgroup
// barriers); tile t of iteration it belongs to sub-tile (t / grid) mod S of
// workgroup t mod grid, so a batch of fewer tiles than CUs still spreads.
//
// A workgroup holds 12 waves (12 / K tiles: 6, 4, 3 for ranks 2, 3, 4),
// three per SIMD.  Measured against the most tiles LDS allows (7 at rank 2,
// 5 at rank 3: 14 and 15 waves) and the fused kernel at rank 4 (16 waves of
// four-wave workgroups), per launch: rank 2 61.5 -> 55.6 us, rank 3 93.0 ->
// 88.3 us, rank 4 141.1 -> 131.2 us at 65,536 instances, and 4-7% less at
// 81,920 and 131,072 (rank 3: 109.4 -> 107.3 and 172.6 -> 165.6 us, where
// five tiles per round come out even); 5 or 3 tiles at ranks 2 and 3 (10 or
// 9 waves) measured slower again.
// ---------------------------------------------------------------------------
// biases of both per-block passes, fwd2 + inv2 matrices, fwd1 + inv1p
// matrices and their biases
constexpr int kWideLdsTables = 2 * 1024 + 2 * 16384 + 2 * 1024 + 2 * 64;
constexpr int kWideWaves = 12;
template <int K>
constexpr int wide_tiles()
{
  return kWideWaves / K;
}
// kWideWaves slabs (K divides 12 for every rank) and the tables.
static_assert(kWideLdsTables + kWideWaves * kTileInst * kStride * 2 <= 163840,
              "the wide kernels' slabs and tables fit the workgroup's LDS");
constexpr int kWidePre = 3;  // matrix-row chunks in flight

// SBITS = 2, 3, 4: s packed (matvec_ntt_fused_packed).  Each wave holds the
// packed bytes of its s_w in the lane layout of the direct first pass
// (pass_m16_packed), loaded with the default cache policy as
// packed_load_lanes does, instead of the int16 rows; the first-pass matrix
// and bias in LDS are the packed digits' (m16_fwd1pk) in place of fwd1's.
// 65,536 instances, against the fused kernel's four-wave workgroups: rank 3
// 84.4 against 88.8 us (3 bits), 82.8 against 87.7 (2 bits); rank 2 50.9
// against 55.4; rank 4 equal (128.1 against 127.5).
template <int K, int SBITS = 0>
__global__ __launch_bounds__(64 * K * wide_tiles<K>()) __attribute__((amdgpu_waves_per_eu(3)))
void k_matvec_wide(const FusedIo io, const DeviceTables *__restrict__ tab)
{
  constexpr int P = kTileInst, S = wide_tiles<K>(), W = K * S;
  __shared__ __align__(16) int16_t s_slab[S][K][P * kStride];
  __shared__ __align__(16) int32_t s_bf2[16][16];
  __shared__ __align__(16) int32_t s_bi2[16][16];
  __shared__ __align__(16) uint64_t s_m2[2][16][2][64];  // [fwd2, inv2][block][limb][lane]
  __shared__ __align__(16) uint64_t s_m1[2][2][64];      // [fwd1, inv1p][limb][lane]
  __shared__ __align__(16) int32_t s_b1[2][16];          // bias_fwd1, bias_inv1p
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int st = w / K, wi = w % K;  // sub-tile, wave within it
  const int lane = threadIdx.x & 63, g = lane >> 4, tt = lane & 15;

  for (int i = threadIdx.x; i < 256; i += 64 * W) {
    s_bf2[i >> 4][i & 15] = tab->bias_fwd2[i >> 4][i & 15];
    s_bi2[i >> 4][i & 15] = tab->bias_inv2[i >> 4][i & 15];
  }
  for (int i = threadIdx.x; i < 16 * 2 * 64; i += 64 * W) {
    s_m2[0][i >> 7][(i >> 6) & 1][i & 63] =
        *reinterpret_cast<const uint64_t *>(tab->m16_fwd2[i >> 7][(i >> 6) & 1][i & 63]);
    s_m2[1][i >> 7][(i >> 6) & 1][i & 63] =
        *reinterpret_cast<const uint64_t *>(tab->m16_inv2[i >> 7][(i >> 6) & 1][i & 63]);
  }
  constexpr int kPk = SBITS ? SBITS - 2 : 0;
  for (int i = threadIdx.x; i < 2 * 64; i += 64 * W) {
    s_m1[0][i >> 6][i & 63] =
        *reinterpret_cast<const uint64_t *>(SBITS ? tab->m16_fwd1pk[kPk][i >> 6][i & 63]
                                                  : tab->m16_fwd1[0][i >> 6][i & 63]);
    s_m1[1][i >> 6][i & 63] = *reinterpret_cast<const uint64_t *>(tab->m16_inv1p[0][i >> 6][i & 63]);
  }
  for (int i = threadIdx.x; i < 16; i += 64 * W) {
    s_b1[0][i] = SBITS ? tab->bias_fwd1pk[0][i] : tab->bias_fwd1[0][i];
    s_b1[1][i] = tab->bias_inv1p[0][i];
  }
  __syncthreads();

  // Every matrix and bias is read from LDS at each pass, keeping registers
  // for the loads in flight.
  auto m1 = [&](int which, int limb) { return (long)s_m1[which][limb][lane]; };
  auto b1 = [&](int which) { return *reinterpret_cast<const v4i *>(&s_b1[which][4 * g]); };
  auto f2lo = [&](int m) { return (long)s_m2[0][m][0][lane]; };
  auto f2hi = [&](int m) { return (long)s_m2[0][m][1][lane]; };
  auto i2lo = [&](int m) { return (long)s_m2[1][m][0][lane]; };
  auto i2hi = [&](int m) { return (long)s_m2[1][m][1][lane]; };
  uint32_t gin[4];
  {
    const uint32_t ga = tab->gamma[2 * (lane & 31)], gb = tab->gamma[2 * (lane & 31) + 1];
    gin[0] = ga;
    gin[1] = (uint32_t)kQ - ga;
    gin[2] = gb;
    gin[3] = (uint32_t)kQ - gb;
  }

  int16_t *const slab = s_slab[st][wi];
  int16_t *const pslab = slab + tt * kStride;
  const size_t ntiles = (io.ninst + P - 1) / P, G = gridDim.x, per_iter = (size_t)S * G;
  auto tile_of = [&](size_t base) { return base + (size_t)st * G + blockIdx.x; };
  auto nvalid_of = [&](size_t tile) {
    const size_t i0 = tile * P;
    return tile < ntiles ? (int)(io.ninst - i0 < (size_t)P ? io.ninst - i0 : P) : 0;
  };
  auto rt = [&](int role, size_t tile, int nvalid) {
    return RoleTile<false>(role_base(io, role), tile * P, nvalid,
                           (int)role_stride<kFusedMatvec, K>(role), lane);
  };
  // s_w of this sub-tile's first tile; later ones are prefetched in phase 3.
  // Packed: lane (g, t) takes the bytes of rows 4g..4g+3 of poly wi of
  // instance t, through a descriptor over the tile's valid instances.
  constexpr int PPB = SBITS ? packed_poly_bytes<SBITS>() : 1;
  constexpr int NDW = SBITS ? packed_lane_dwords<SBITS>() : 4;
  v4i sx[SBITS ? 1 : P / 2];
  uint32_t pv[NDW];
  auto load_x = [&](size_t t) {
    const int nv = nvalid_of(t);
    if constexpr (SBITS == 0) {
      const auto xr = rt(rX, t, nv);
#pragma unroll
      for (int it = 0; it < P / 2; it++)
        sx[it] = xr.load(it, wi);
    } else {
      const __amdgpu_buffer_rsrc_t xr =
          span_rsrc(nv ? static_cast<const uint8_t *>(io.x) + t * P * K * PPB : nullptr,
                    (uint32_t)nv * K * PPB);
#pragma unroll
      for (int i = 0; i < NDW / 4; i++) {
        const v4i x = __builtin_bit_cast(
            v4i, __builtin_amdgcn_raw_buffer_load_b128(xr, tt * K * PPB + g * NDW * 4,
                                                       wi * PPB + i * 16, kBufferKeep));
        pv[4 * i] = (uint32_t)x[0];
        pv[4 * i + 1] = (uint32_t)x[1];
        pv[4 * i + 2] = (uint32_t)x[2];
        pv[4 * i + 3] = (uint32_t)x[3];
      }
    }
  };
  load_x(tile_of(0));
  vm_drain();  // the loop's waits for s_w then count the previous tile's stores

  for (size_t base = 0; base < ntiles; base += per_iter) {
    const size_t tile = tile_of(base);
    const int nvalid = nvalid_of(tile);  // 0: an idle sub-tile (it still meets the barriers)
    // 1. s_w -> slab, the row's first chunks in flight, forward transform.
    const auto mr = rt(rMat, tile, nvalid);
    v4i a[kWidePre][K];
    wave_sync();
    if constexpr (SBITS == 0) {
#pragma unroll
      for (int it = 0; it < P / 2; it++) {
        const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
        *reinterpret_cast<v4i *>(slab + pp * kStride + ch * 8) = sx[it];
      }
    }
#pragma unroll
    for (int c = 0; c < kWidePre; c++)
#pragma unroll
      for (int j = 0; j < K; j++)
        a[c][j] = mr.load(c, wi * K + j);
    if (nvalid) {
      if constexpr (SBITS == 0) {
        wave_sync();
        pass_m16<kFormInter>(pslab, g, m1(0, 0), m1(0, 1), b1(0));
      } else {
        pass_m16_packed<SBITS>(pslab, g, pv, m1(0, 0), m1(0, 1), b1(0));
      }
      wave_sync();
      pass_c16<kFormInter>(pslab, g, f2lo, f2hi, s_bf2);
    }
    __syncthreads();

    // 2. Row wi: kWidePre chunks of the matrix row in flight.
    uint32_t res[P / 2][4];
#pragma unroll
    for (int it = 0; it < P / 2 && nvalid; it++) {
      const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
      v4i *cur = a[it % kWidePre];
      int r0[4] = {0, 0, 0, 0}, r1[4] = {0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < K; j++) {
        const v4i b = *reinterpret_cast<const v4i *>(s_slab[st][j] + pp * kStride + ch * 8);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t bu = (uint32_t)b[u];
          const uint32_t x = __builtin_amdgcn_perm(red32((bu >> 16) * gin[u]), bu, 0x05040100u);
          const uint32_t y = __builtin_amdgcn_alignbit(bu, bu, 16);
          r0[u] = dot2((uint32_t)cur[j][u], x, r0[u]);
          r1[u] = dot2((uint32_t)cur[j][u], y, r1[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        res[it][u] =
            pack2(red32((uint32_t)r0[u] + kSumOffset), red32((uint32_t)r1[u] + kSumOffset));
      if (it + kWidePre < P / 2)
#pragma unroll
        for (int j = 0; j < K; j++)
          cur[j] = mr.load(it + kWidePre, wi * K + j);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();  // every wave is done reading the slabs

    // 3. The row into the slab, the next tile's s_w in flight, inverse, t.
#pragma unroll
    for (int it = 0; it < P / 2; it++) {
      const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
      *reinterpret_cast<v4i *>(slab + pp * kStride + ch * 8) =
          v4i{(int)res[it][0], (int)res[it][1], (int)res[it][2], (int)res[it][3]};
    }
    load_x(tile + per_iter);
    if (nvalid) {
      wave_sync();
      pass_c16<kFormInter>(pslab, g, i2lo, i2hi, s_bi2);
      wave_sync();
      pass_m16<kFormCentred>(pslab, g, m1(1, 0), m1(1, 1), b1(1));
    }
    // Also for an idle sub-tile, whose descriptor has size 0 and drops the
    // stores: every path into the next iteration then has the same
    // operations in flight, and the wait for s_w stays behind the stores.
    wave_sync();
    role_out<P>(slab, rt(rOut0, tile, nvalid), wi, lane);
  }
}

constexpr int kDecryptWaves = 4;
constexpr int kDecryptWideWaves = 16;

// Decrypt, one wave per tile: w = reduce(v - invntt(sum_j s_hat_j o
// ntt(u_j))).  Decrypt has one output row, so a wave forms it alone: for
// each j it stages u_j of its 16 instances into its slab, forward-transforms
// it and adds s_hat_j o u_hat_j (s_hat_j streamed from HBM) into the row it
// holds in registers, packed; then it inverse-transforms the row in the slab
// and writes w.  The waves of a workgroup share only the bias tables (and,
// wide, the forward matrices), so the tile loop has no workgroup barrier.
// gamma * b1 is formed where it is used, as in the fused kernel's rows.
//
// The row's partial sums are reduced per j to [0, 2q) and added as packed
// int16 pairs (below 8q for K <= 4: no carry between the halves); the
// inverse's first pass takes inputs below 8q, as the fused kernel's rows.
//
// Two shapes, both four waves per SIMD (at most 128 VGPRs):
//   - WIDE (contiguous buffers): one workgroup of 16 waves per CU, holding
//     the forward transform's per-block matrices in LDS (16 KB, shared by
//     the 16 waves; 153,600 B of LDS in all).  The forward transform then
//     issues no vector memory load, so s_hat_j's loads go out before it and
//     are in flight for its whole length: vector memory loads complete in
//     order, and a matrix fetched from memory inside the transform would wait
//     for them.  Waves stride over the tiles (the grid is one workgroup per
//     CU).
//   - four-wave workgroups fetching the matrices from the (cache-resident)
//     table, s_hat_j loaded after the forward transform: the table-driven
//     variant, whose 64-bit pointer rows take the LDS; it runs at three
//     waves per SIMD rather than spill.
template <int K, bool TABLE, bool WIDE>
__global__ __launch_bounds__(64 * (WIDE ? kDecryptWideWaves : kDecryptWaves))
__attribute__((amdgpu_waves_per_eu(TABLE ? 3 : 4))) void k_decrypt(
    const FusedIo io, const DeviceTables *__restrict__ tab)
{
  static_assert(!(TABLE && WIDE), "the table-driven variant is four-wave");
  constexpr int P = kTileInst, W = WIDE ? kDecryptWideWaves : kDecryptWaves;
  __shared__ const int16_t *s_ip[TABLE ? W : 1][TABLE ? P : 1][8];
  __shared__ __align__(16) int16_t s_slab[W][P * kStride];
  __shared__ __align__(16) int32_t s_bf2[16][16];
  __shared__ __align__(16) int32_t s_bi2[16][16];
  // Wide: the forward per-block matrices and both first-pass matrices,
  // [block][limb][lane] (fwd2), [limb][lane] (fwd1, inv1p).
  __shared__ __align__(16) uint64_t s_f2[WIDE ? 16 : 1][2][64];
  __shared__ __align__(16) uint64_t s_p1[WIDE ? 2 : 1][2][64];
  __shared__ __align__(16) int32_t s_b1[WIDE ? 2 : 1][16];  // bias_fwd1, bias_inv1p
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, g = lane >> 4, tt = lane & 15;

  for (int i = threadIdx.x; i < 256; i += 64 * W) {
    s_bf2[i >> 4][i & 15] = tab->bias_fwd2[i >> 4][i & 15];
    s_bi2[i >> 4][i & 15] = tab->bias_inv2[i >> 4][i & 15];
  }
  if constexpr (WIDE) {
    for (int i = threadIdx.x; i < 16 * 2 * 64; i += 64 * W)
      s_f2[i >> 7][(i >> 6) & 1][i & 63] =
          *reinterpret_cast<const uint64_t *>(tab->m16_fwd2[i >> 7][(i >> 6) & 1][i & 63]);
    for (int i = threadIdx.x; i < 2 * 64; i += 64 * W) {
      s_p1[0][i >> 6][i & 63] = *reinterpret_cast<const uint64_t *>(tab->m16_fwd1[0][i >> 6][i & 63]);
      s_p1[1][i >> 6][i & 63] = *reinterpret_cast<const uint64_t *>(tab->m16_inv1p[0][i >> 6][i & 63]);
    }
    for (int i = threadIdx.x; i < 16; i += 64 * W) {
      s_b1[0][i] = tab->bias_fwd1[0][i];
      s_b1[1][i] = tab->bias_inv1p[0][i];
    }
  }
  __syncthreads();

  // First-pass matrices: held in registers, or (wide) read from LDS at each
  // pass, which keeps their 8 VGPRs free for s_hat_j in flight.
  const long f1lo_r = WIDE ? 0 : ld_a(tab->m16_fwd1[0], 0, lane);
  const long f1hi_r = WIDE ? 0 : ld_a(tab->m16_fwd1[0], 1, lane);
  const long i1lo_r = WIDE ? 0 : ld_a(tab->m16_inv1p[0], 0, lane);
  const long i1hi_r = WIDE ? 0 : ld_a(tab->m16_inv1p[0], 1, lane);
  auto p1 = [&](int which, int limb, long reg) {
    if constexpr (WIDE)
      return (long)s_p1[which][limb][lane];
    else
      return reg;
  };
  auto b1 = [&](int which, const v4i &reg) {
    if constexpr (WIDE)
      return *reinterpret_cast<const v4i *>(&s_b1[which][4 * g]);
    else
      return reg;
  };
  const v4i fb1 = *reinterpret_cast<const v4i *>(&tab->bias_fwd1[0][4 * g]);
  const v4i ib1 = *reinterpret_cast<const v4i *>(&tab->bias_inv1p[0][4 * g]);
  const __amdgpu_buffer_rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<DeviceTables *>(tab), 0, (int)sizeof(DeviceTables), kBufferRsrcWord3);
  const int loff = lane * 8;
  auto mat = [&](int off) {
    return __builtin_bit_cast(long, __builtin_amdgcn_raw_buffer_load_b64(trs, loff, off, 0));
  };
  constexpr int kF2 = (int)offsetof(DeviceTables, m16_fwd2), kI2 = (int)offsetof(DeviceTables, m16_inv2);
  auto f2lo = [&](int m) {
    if constexpr (WIDE)
      return (long)s_f2[m][0][lane];
    else
      return mat(kF2 + m * 1024);
  };
  auto f2hi = [&](int m) {
    if constexpr (WIDE)
      return (long)s_f2[m][1][lane];
    else
      return mat(kF2 + m * 1024 + 512);
  };
  auto i2lo = [&](int m) { return mat(kI2 + m * 1024); };
  auto i2hi = [&](int m) { return mat(kI2 + m * 1024 + 512); };
  // Lane (., ch) holds coefficients 8ch .. 8ch+7 of a poly: pairs u = 0..3,
  // pair u in group 2ch + u/2 with gamma for even u and q - gamma for odd u.
  uint32_t gam[4];
  {
    const uint32_t ga = tab->gamma[2 * (lane & 31)], gb = tab->gamma[2 * (lane & 31) + 1];
    gam[0] = ga;
    gam[1] = (uint32_t)kQ - ga;
    gam[2] = gb;
    gam[3] = (uint32_t)kQ - gb;
  }

  int16_t *const slab = s_slab[w];
  int16_t *const pslab = slab + tt * kStride;
  size_t i0 = 0;
  int nvalid = 0;
  auto role = [&](int rl) {
    if constexpr (TABLE)
      return RoleTile<true>{s_ip[w], rl, lane};
    else
      return RoleTile<false>(role_base(io, rl), i0, nvalid,
                             (int)role_stride<kFusedDecrypt, K>(rl), lane);
  };

  // Tile t goes to workgroup t mod grid, wave (t / grid) mod W: a batch of
  // fewer tiles than wave slots spreads over every CU.
  const size_t ntiles = (io.ninst + P - 1) / P, nwaves = (size_t)gridDim.x * W;
  for (size_t tile = WIDE ? (size_t)w * gridDim.x + blockIdx.x : (size_t)blockIdx.x * W + w;
       tile < ntiles; tile += nwaves) {
    i0 = tile * P;
    nvalid = (int)(io.ninst - i0 < (size_t)P ? io.ninst - i0 : P);
    if constexpr (TABLE) {
      // The wave's own rows of pointers; its reads of the previous tile's
      // rows were issued before these writes.
      wave_sync();
      for (int idx = lane; idx < P * 8; idx += 64) {
        const int pp = idx >> 3, rl = idx & 7;
        s_ip[w][pp][rl] = pp < nvalid
                              ? reinterpret_cast<const int16_t *const *>(io.inst + i0 + pp)[rl]
                              : nullptr;
      }
      wave_sync();
    }

    // Software-pipelined over the tile's HBM reads: u_{j+1} is in flight
    // during the product with s_hat_j, whose loads all go out at once.
    const auto um = role(rX);
    const auto sm = role(rMat);
    uint32_t acc[P / 2][4];
    v4i un[P / 2];
#pragma unroll
    for (int it = 0; it < P / 2; it++) {
      acc[it][0] = acc[it][1] = acc[it][2] = acc[it][3] = 0;
      un[it] = um.load(it, 0);
    }
    // One j: u_hat_j = ntt(u_j) (canonical, in the slab), then acc +=
    // s_hat_j o u_hat_j with chunk `it` of u_{j+1} loaded into the registers
    // chunk `it` of s_hat_j frees.  The last j is a separate instance that
    // loads nothing: a conditional load would keep the old chunk values live
    // across the forward transform.
    auto step = [&](int j, auto last_c) {
      constexpr bool kLast = decltype(last_c)::value;
      wave_sync();
#pragma unroll
      for (int it = 0; it < P / 2; it++) {
        const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
        *reinterpret_cast<v4i *>(slab + pp * kStride + ch * 8) = un[it];
      }
      v4i sn[P / 2];
      if constexpr (WIDE) {
        // In flight during the whole forward transform, which loads nothing.
#pragma unroll
        for (int it = 0; it < P / 2; it++)
          sn[it] = sm.load(it, j);
      }
      wave_sync();
      pass_m16<kFormInter>(pslab, g, p1(0, 0, f1lo_r), p1(0, 1, f1hi_r), b1(0, fb1));
      wave_sync();
      pass_c16<kFormInter>(pslab, g, f2lo, f2hi, s_bf2);
      wave_sync();
      if constexpr (!WIDE) {
#pragma unroll
        for (int it = 0; it < P / 2; it++)
          sn[it] = sm.load(it, j);
      }
#pragma unroll
      for (int it = 0; it < P / 2; it++) {
        const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
        const v4i b = *reinterpret_cast<const v4i *>(slab + pp * kStride + ch * 8);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t bu = (uint32_t)b[u];
          const uint32_t gb1 = red32((bu >> 16) * gam[u]);
          const uint32_t x = __builtin_amdgcn_perm(gb1, bu, 0x05040100u);
          const uint32_t y = __builtin_amdgcn_alignbit(bu, bu, 16);
          const uint32_t r = pack2(red32((uint32_t)dot2((uint32_t)sn[it][u], x, 0) + kSumOffset),
                                   red32((uint32_t)dot2((uint32_t)sn[it][u], y, 0) + kSumOffset));
          acc[it][u] += r;
        }
        if constexpr (!kLast)
          un[it] = um.load(it, j + 1);
        else if constexpr (WIDE)
          un[it] = role(rAdd0).load(it, 0);  // v, for the epilogue
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    // Wide: the products run at wave priority 1 and the inverse and the
    // epilogue at 0, so a SIMD issues first for the waves still streaming
    // u_j and s_hat_j in, and the waves near their tile's end fill the gaps.
    // 65,536 instances, one process: 49.6-50.0 against 53.7-54.0 us at rank
    // 3, 39.5 against 42.1 at rank 2, 60.7 against 64.5 at rank 4.  (Every
    // phase at its own level, 3 down to 0, took rank 3 to 66 us: waves kept
    // in step load and compute at the same moments.)
    if constexpr (WIDE)
      __builtin_amdgcn_s_setprio(1);
#pragma unroll 1
    for (int j = 0; j < K - 1; j++)
      step(j, std::false_type{});
    step(K - 1, std::true_type{});
    if constexpr (WIDE)
      __builtin_amdgcn_s_setprio(0);

    // The row into the slab (its reads above are done), inverse-transformed.
    wave_sync();
#pragma unroll
    for (int it = 0; it < P / 2; it++) {
      const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
      *reinterpret_cast<v4i *>(slab + pp * kStride + ch * 8) =
          v4i{(int)acc[it][0], (int)acc[it][1], (int)acc[it][2], (int)acc[it][3]};
    }
    wave_sync();
    pass_c16<kFormInter>(pslab, g, i2lo, i2hi, s_bi2);
    wave_sync();
    // w = reduce(v - row); every load of v goes out before the first store
    // (w may alias v).  Wide, they went out during the last product, into
    // the registers s_hat's chunks freed.
    const auto va = role(rAdd0);
    const auto oa = role(rOut0);
    pass_m16<kFormCentred>(pslab, g, p1(1, 0, i1lo_r), p1(1, 1, i1hi_r), b1(1, ib1));
    wave_sync();
    const v4i zero = {0, 0, 0, 0};
    if constexpr (!WIDE) {
#pragma unroll
      for (int it = 0; it < P / 2; it++)
        un[it] = va.load(it, 0);
    }
#pragma unroll
    for (int it = 0; it < P / 2; it++) {
      const int q = it * 64 + lane, pp = q >> 5, ch = q & 31;
      const v4i x = *reinterpret_cast<const v4i *>(slab + pp * kStride + ch * 8);
      oa.store(it, 0, centred_sum<-1>(x, un[it], zero));
    }
  }
}

}  // namespace

// Grids of up to 4x the resident workgroups: the dispatcher then hands tiles
// to whichever slot frees up, where a resident-sized grid left a tail of
// workgroups on a last extra tile.  65,536 rank-3 instances: keygen 128 ->
// 125 us, encryption 147 -> 144 us, matvec unchanged; 262,144: decryption
// 243 -> 238-241 us.
constexpr int kFusedOversub = 4;

int launch_fused(int k, int mode, int s_bits, const FusedIo &io, const DeviceTables *tab, int ncu,
                 const LaunchSink &stream)
{
  const size_t ntiles = (io.ninst + kTileInst - 1) / kTileInst;
  auto go = [&](auto kernel, int waves) {
    const size_t cap =
        (size_t)ncu * resident_per_cu((const void *)kernel, 64 * waves) * kFusedOversub;
    const dim3 grid((unsigned)(ntiles < cap ? ntiles : cap));
    return sink_launch(stream, kernel, grid, dim3(64 * waves), io, tab);
  };
  auto with_k = [&](auto mode_c, auto table_c) {
    constexpr int M = decltype(mode_c)::value;
    constexpr bool T = decltype(table_c)::value;
    switch (k) {
      case 2: return go(k_fused<2, M, T>, fused_waves<M, 2>());
      case 3: return go(k_fused<3, M, T>, fused_waves<M, 3>());
      case 4: return go(k_fused<4, M, T>, fused_waves<M, 4>());
      default: return -1;
    }
  };
  using std::integral_constant;
  auto by_table = [&](auto mode_c) {
    return io.inst ? with_k(mode_c, std::true_type{}) : with_k(mode_c, std::false_type{});
  };
  if (io.inst && s_bits)
    return -1;
  // Wide matvec: one workgroup per CU (its LDS), persistent over the tiles.
  auto wide = [&](auto kernel, int waves) {
    const size_t grid = ntiles < (size_t)ncu ? ntiles : (size_t)ncu;
    return sink_launch(stream, kernel, dim3((unsigned)grid), dim3(64 * waves), io, tab);
  };
  switch (mode) {
    case kFusedMatvec: {
      // Contiguous batches (int16 or packed s): the wide kernel; plans: the
      // table-driven fused kernel.
      if (!io.inst) {
        auto wide_k = [&](auto bits_c) {
          constexpr int B = decltype(bits_c)::value;
          switch (k) {
            case 2: return wide(k_matvec_wide<2, B>, 2 * wide_tiles<2>());
            case 3: return wide(k_matvec_wide<3, B>, 3 * wide_tiles<3>());
            case 4: return wide(k_matvec_wide<4, B>, 4 * wide_tiles<4>());
            default: return -1;
          }
        };
        switch (s_bits) {
          case 0: return wide_k(integral_constant<int, 0>{});
          case 2: return wide_k(integral_constant<int, 2>{});
          case 3: return wide_k(integral_constant<int, 3>{});
          case 4: return wide_k(integral_constant<int, 4>{});
          default: return -1;
        }
      }
      return with_k(integral_constant<int, kFusedMatvec>{}, std::true_type{});
    }
    case kFusedKeygen:
      return s_bits ? -1 : by_table(integral_constant<int, kFusedKeygen>{});
    case kFusedEncrypt:
      return s_bits ? -1 : by_table(integral_constant<int, kFusedEncrypt>{});
    case kFusedDecrypt: {
      if (s_bits)
        return -1;
      // Wide: one workgroup of 16 waves per CU, the waves striding over the
      // tiles.  Table-driven: four-wave workgroups, up to kFusedOversub x the
      // resident ones.
      auto dec = [&](auto kernel, int waves, int oversub) {
        const size_t cap =
            (size_t)ncu * resident_per_cu((const void *)kernel, 64 * waves) * oversub;
        const size_t wgs = waves == kDecryptWideWaves ? ntiles : (ntiles + waves - 1) / waves;
        return sink_launch(stream, kernel, dim3((unsigned)(wgs < cap ? wgs : cap)),
                           dim3(64 * waves), io, tab);
      };
      constexpr int kW = kDecryptWideWaves, kT = kDecryptWaves;
      switch (k * 2 + (io.inst ? 1 : 0)) {
        case 4: return dec(k_decrypt<2, false, true>, kW, 1);
        case 5: return dec(k_decrypt<2, true, false>, kT, kFusedOversub);
        case 6: return dec(k_decrypt<3, false, true>, kW, 1);
        case 7: return dec(k_decrypt<3, true, false>, kT, kFusedOversub);
        case 8: return dec(k_decrypt<4, false, true>, kW, 1);
        case 9: return dec(k_decrypt<4, true, false>, kT, kFusedOversub);
        default: return -1;
      }
    }
    default:
      return -1;
  }
}

}  // namespace tk
