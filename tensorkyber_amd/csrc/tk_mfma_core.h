// Device building blocks of the INT8-MFMA transform, shared by the batch,
// ragged and fused kernels: tile staging between HBM and a per-wave LDS
// slab, and the two contraction passes for the 16x16x32 and 32x32x32 tile
// shapes.  See tk_mfma.hip for the formulation.
#pragma once

#include "tk_device.h"

namespace tk {

constexpr int kStride = 264;  // int16 per poly slot in LDS (512 + 16 bytes)
// MFMA pairs in flight per pass step: four columns or blocks per step keep 8
// MFMAs in flight with 64-bit LDS accesses.
constexpr int kPassGroup = 4;

// ---------------------------------------------------------------------------
// Tile staging: P polys, coalesced 16-byte global accesses, one LDS slab
// (poly pp at pp * kStride).  Chunk `it` of a tile is polys 2*it and 2*it+1,
// lanes 0-31 and 32-63, 16 bytes per lane.
// ---------------------------------------------------------------------------
// Tile staging through a buffer descriptor over the tile's own bytes (P
// consecutive polys, fewer in the last tile of a batch).  Accesses past the
// descriptor's size read zero (loads) or are dropped (stores), so a partial
// tile needs no per-lane checks: the loads and the stores go out as one
// unconditional batch each.
//
// Stores take their whole offset in the VGPR operand and a zero scalar
// offset.  A 16-byte store whose data VGPRs are overwritten by the very next
// VALU instruction writes the new values on gfx950; the compiler inserts the
// wait state that prevents this only for stores without a scalar offset
// register (with one it assumes the hardware needs none).  Loads read no
// data VGPRs and keep the cheaper scalar offset (the chunk offset): on
// gfx950 the range check of these loads (raw buffer, VGPR offset enabled)
// covers the scalar offset too.  Measured: loads of a full tile whose
// descriptor size was reduced by their scalar offset returned zeros from
// the second half of the tile, and batches ending at the last byte of their
// allocation read nothing past it (tests/test_gpu_bounds.py).
constexpr int kBufferRsrcWord3 = 0x00020000;  // gfx9 buffer descriptor: 32-bit data format
constexpr int kBufferNT = 2;                  // cache policy: nontemporal
constexpr int kBufferKeep = 0;                // cache policy: default (allocates in the caches)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t span_rsrc(const void *base, uint32_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes,
                                           kBufferRsrcWord3);
}

// Groups of eight 16-byte accesses per lane (one 16-poly tile, or half of a
// 32-poly one), with a scheduling barrier between groups so that at most one
// group's registers are live.
template <int P>
__device__ __forceinline__ void tile_in_rsrc(int16_t *L, __amdgpu_buffer_rsrc_t r, int lane)
{
#pragma unroll
  for (int g0 = 0; g0 < P / 2; g0 += 8) {
    v4i v[8];
#pragma unroll
    for (int i = 0; i < 8; i++)
      v[i] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(
                                         r, lane * 16, (g0 + i) * 1024, kBufferNT));
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int q = (g0 + i) * 64 + lane, pp = q >> 5, ch = q & 31;
      *reinterpret_cast<v4i *>(L + pp * kStride + ch * 8) = v[i];
    }
    if (g0 + 8 < P / 2)
      __builtin_amdgcn_sched_barrier(0);
  }
}

// POL: the stores' cache policy (kBufferNT, or kBufferKeep to leave the
// tile in the caches for the batch's next pass).
template <int P, int POL = kBufferNT>
__device__ __forceinline__ void tile_out_rsrc(const int16_t *L, __amdgpu_buffer_rsrc_t r, int lane)
{
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int g0 = 0; g0 < P / 2; g0 += 8) {
    v4i v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int q = (g0 + i) * 64 + lane, pp = q >> 5, ch = q & 31;
      v[i] = *reinterpret_cast<const v4i *>(L + pp * kStride + ch * 8);
    }
#pragma unroll
    for (int i = 0; i < 8; i++)
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, v[i]), r,
                                             lane * 16 + (g0 + i) * 1024, 0, POL);
    if (g0 + 8 < P / 2)
      __builtin_amdgcn_sched_barrier(0);
  }
}

// Write a 16-poly tile from the slab (rout) and load the next tile into it
// (rin).  The next tile's loads go out before the stores, so waiting for
// them, before their LDS writes, does not wait for these stores as well
// (vector memory operations complete in issue order); the stores complete
// during the next tile's passes.  The stores' data goes through registers
// two chunks at a time (a scheduling barrier per pair), so the swap holds 8
// registers beyond the loads' 32 and the kernel keeps 128 VGPRs.  POL is the
// stores' cache policy, as for tile_out_rsrc.
template <int POL = kBufferNT>
__device__ __forceinline__ void tile_swap_rsrc16(int16_t *L, __amdgpu_buffer_rsrc_t rout,
                                                 __amdgpu_buffer_rsrc_t rin, int lane)
{
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  v4i n[8];
#pragma unroll
  for (int i = 0; i < 8; i++)
    n[i] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rin, lane * 16, i * 1024,
                                                                         kBufferNT));
#pragma unroll
  for (int h = 0; h < 8; h += 2) {
    v4i o[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int q = (h + i) * 64 + lane, pp = q >> 5, ch = q & 31;
      o[i] = *reinterpret_cast<const v4i *>(L + pp * kStride + ch * 8);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, o[i]), rout,
                                             lane * 16 + (h + i) * 1024, 0, POL);
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int q = i * 64 + lane, pp = q >> 5, ch = q & 31;
    *reinterpret_cast<v4i *>(L + pp * kStride + ch * 8) = n[i];
  }
}

// ---------------------------------------------------------------------------
// Packed small-coefficient input, the format of the reference's pack_small
// (ntt_bench.cu:749-757): BITS-bit two's-complement fields, 8/BITS fields per
// byte, coefficient i of a poly in byte i/(8/BITS) at bit BITS*(i%(8/BITS)).
// A poly takes 256/(8/BITS) bytes: 64 for BITS = 2, 128 for BITS = 3 or 4.
// ---------------------------------------------------------------------------
template <int BITS>
constexpr int packed_poly_bytes()
{
  return kN / (8 / BITS);
}

// One packed dword (coefficients c0 .. c0 + 4*(8/BITS) - 1) -> int16 at dst
// (16-byte aligned, contiguous coefficients), sign-extending each field.
template <int BITS>
__device__ __forceinline__ void unpack_dword(uint32_t raw, int16_t *dst)
{
  constexpr int PER = 8 / BITS;  // fields per byte
  constexpr int CPD = 4 * PER;   // coefficients per dword: 16 or 8
  uint32_t pr[CPD / 2];
#pragma unroll
  for (int j = 0; j < CPD / 2; j++) {
    const int i0 = 2 * j, i1 = 2 * j + 1;
    const int lo = __builtin_amdgcn_sbfe((int)raw, 8 * (i0 / PER) + BITS * (i0 % PER), BITS);
    const int hi = __builtin_amdgcn_sbfe((int)raw, 8 * (i1 / PER) + BITS * (i1 % PER), BITS);
    pr[j] = pack2((uint32_t)lo, (uint32_t)hi);
  }
#pragma unroll
  for (int v = 0; v < CPD / 8; v++)
    *reinterpret_cast<v4i *>(dst + 8 * v) =
        v4i{(int)pr[4 * v], (int)pr[4 * v + 1], (int)pr[4 * v + 2], (int)pr[4 * v + 3]};
}

// Packed tile staging, split so the loads of the next tile can be in flight
// during the current one: packed_load reads P polys of
// packed_poly_bytes<BITS>() each, one dword per lane (consecutive lanes take
// consecutive dwords: coalesced), into v[P * bytes / 256]; packed_park
// unpacks them into the int16 slab (neighbouring lanes write adjacent 16 or
// 32 bytes: free of LDS bank conflicts).  `src` yields the packed address of
// tile poly pp (nullptr = beyond the batch: zeros).
template <int P, int BITS>
constexpr int packed_regs()
{
  return P * packed_poly_bytes<BITS>() / 4 / 64;
}

template <int P, int BITS, class Addr>
__device__ __forceinline__ void packed_load(uint32_t *v, Addr src, int lane)
{
  constexpr int DPP = packed_poly_bytes<BITS>() / 4;  // dwords per poly
  static_assert(P * DPP % 64 == 0, "whole dwords per lane");
#pragma unroll
  for (int it = 0; it < packed_regs<P, BITS>(); it++) {
    const int q = it * 64 + lane, pp = q / DPP, wd = q % DPP;
    const uint8_t *p = src(pp);
    v[it] = p ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p) + wd) : 0u;
  }
}

template <int P, int BITS>
__device__ __forceinline__ void packed_park(int16_t *L, const uint32_t *v, int lane)
{
  constexpr int DPP = packed_poly_bytes<BITS>() / 4;
  constexpr int CPD = 4 * (8 / BITS);  // coefficients per dword
#pragma unroll
  for (int it = 0; it < packed_regs<P, BITS>(); it++) {
    const int q = it * 64 + lane, pp = q / DPP, wd = q % DPP;
    unpack_dword<BITS>(v[it], L + pp * kStride + wd * CPD);
  }
}

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

// Final form of two packed values in [0, 2q): kFormInter keeps them,
// kFormCanon maps to [0, q), kFormCentred (whose bias carried +kCentre) to
// [-(q-1)/2, (q-1)/2]; packed 16-bit arithmetic, one pair per instruction.
template <int MODE>
__device__ __forceinline__ uint32_t finish2(uint32_t w)
{
  if (MODE == kFormInter)
    return w;
  us2 v = __builtin_bit_cast(us2, w);
  const us2 qq = {(unsigned short)kQ, (unsigned short)kQ};
  v = __builtin_elementwise_min(v, (us2)(v - qq));  // u16 wrap: v - q is large when v < q
  if (MODE == kFormCentred) {
    const us2 cc = {(unsigned short)kCentre, (unsigned short)kCentre};
    v = v - cc;
  }
  return __builtin_bit_cast(uint32_t, v);
}

// Four rows of four int16 each (R[r] = row r as two packed pairs: columns
// (0,1) and (2,3)) -> the 8-byte MFMA operand of each column: [x0' of rows
// 0..3 | x1' of rows 0..3], i.e. the digits of a column with its K index
// running over the rows.  A byte transpose in registers (v_perm_b32).
__device__ __forceinline__ void col_digits4(const uint2 R[4], long b[4])
{
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t d0 = h ? R[0].y : R[0].x, d1 = h ? R[1].y : R[1].x;
    const uint32_t d2 = h ? R[2].y : R[2].x, d3 = h ? R[3].y : R[3].x;
    const uint32_t l01 = __builtin_amdgcn_perm(d1, d0, 0x06020400u);
    const uint32_t l23 = __builtin_amdgcn_perm(d3, d2, 0x06020400u);
    const uint32_t h01 = __builtin_amdgcn_perm(d1, d0, 0x07030501u);
    const uint32_t h23 = __builtin_amdgcn_perm(d3, d2, 0x07030501u);
    const uint32_t lo0 = __builtin_amdgcn_perm(l23, l01, 0x05040100u) ^ 0x80808080u;
    const uint32_t lo1 = __builtin_amdgcn_perm(l23, l01, 0x07060302u) ^ 0x80808080u;
    const uint32_t hi0 = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
    const uint32_t hi1 = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
    b[2 * h] = (long)(((uint64_t)hi0 << 32) | lo0);
    b[2 * h + 1] = (long)(((uint64_t)hi1 << 32) | lo1);
  }
}

__device__ __forceinline__ uint2 ld2(const int16_t *p) { return *reinterpret_cast<const uint2 *>(p); }
__device__ __forceinline__ void st2(int16_t *p, uint32_t a, uint32_t b)
{
  *reinterpret_cast<uint2 *>(p) = make_uint2(a, b);
}

// ---------------------------------------------------------------------------
// 16x16x32 passes: lane l = 16g + t works on poly t of the tile; its K group
// and its output rows are the index set {4g, .., 4g+3}.  Each pass works on
// four columns / blocks at a time so eight independent MFMAs are in flight
// before their results are read.
// ---------------------------------------------------------------------------

// Contraction over m for each column c (one matrix for all columns).  G
// columns (2 or 4) per group: G pairs of MFMAs are in flight at once.
template <int MODE, int G = kPassGroup>
__device__ __forceinline__ void pass_m16(int16_t *P, int g, long alo, long ahi, v4i bias)
{
  static_assert(G == 2 || G == 4, "group of 2 or 4 columns");
  const v4i zero = {0, 0, 0, 0};
  int16_t *rows = P + 64 * g;  // row 4g; row 4g + r at +16r
#pragma unroll
  for (int cq = 0; cq < 4; cq++) {
    uint2 R[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
      R[r] = ld2(rows + 16 * r + 4 * cq);
    long b[4];
    col_digits4(R, b);
#pragma unroll
    for (int h = 0; h < 4; h += G) {
      v4i lo[G], hi[G];
#pragma unroll
      for (int u = 0; u < G; u++)
        lo[u] = mfma16(alo, b[h + u], bias);
#pragma unroll
      for (int u = 0; u < G; u++)
        hi[u] = mfma16(ahi, b[h + u], zero);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        uint32_t v[G];
#pragma unroll
        for (int u = 0; u < G; u++)
          v[u] = reduce_2q(lo[u][r], hi[u][r]);
        if (G == 4)
          st2(rows + 16 * r + 4 * cq, finish2<MODE>(pack2(v[0], v[1])),
              finish2<MODE>(pack2(v[G - 2], v[G - 1])));
        else
          *reinterpret_cast<uint32_t *>(rows + 16 * r + 4 * cq + h) =
              finish2<MODE>(pack2(v[0], v[G - 1]));
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // one column quad's operands live at a time
  }
}

// First pass straight from packed small-coefficient input (MFMA16, forward
// layers 1-4).  Lane (g, t) holds, in v, the packed bytes of coefficients
// 64g .. 64g+63 of poly t -- rows m = 4g .. 4g+3 of x[m][c] -- as loaded by
// packed_load_lanes: 4 dwords (one row each) for 2 bits, 8 (two per row)
// for 3 or 4 bits.  The data digits are the unsigned field f and its sign
// bit s (x = f - 2^BITS s; operands m16_fwd1pk): for a column c the four
// rows' fields sit in one byte position of the four row words, so a byte
// gather and a shift-and-mask give each digit of four rows at once -- no
// unpacking to int16, no LDS round trip and no byte transpose.  Results go
// to the slab as pass_m16 writes them.
template <int BITS>
constexpr int packed_lane_dwords()
{
  return packed_poly_bytes<BITS>() / 16;
}

// The loads use the default cache policy.  At 3 and 4 bits a lane's 32
// bytes take two 16-byte loads, each covering every other 16 bytes of the
// tile, so both touch every line; nontemporal, about 2% of the lines were
// fetched twice and the 2^20-poly forward took 129-134 us, against 115-117
// with the line kept in L2 from the first load to the second.  At 2 bits the
// default policy measured 1-3% faster as well (114-116 against 119 us).
template <int BITS, class Addr>
__device__ __forceinline__ void packed_load_lanes(uint32_t *v, Addr src, int lane)
{
  constexpr int NDW = packed_lane_dwords<BITS>();
  const int g = lane >> 4, t = lane & 15;
  const uint8_t *p = src(t);
#pragma unroll
  for (int i = 0; i < NDW / 4; i++) {
    const v4i x = p ? *(reinterpret_cast<const v4i *>(p + g * NDW * 4) + i) : v4i{0, 0, 0, 0};
    v[4 * i] = (uint32_t)x[0];
    v[4 * i + 1] = (uint32_t)x[1];
    v[4 * i + 2] = (uint32_t)x[2];
    v[4 * i + 3] = (uint32_t)x[3];
  }
}

// Byte j of each of d0..d3, as bytes 0..3.
__device__ __forceinline__ uint32_t gather_byte(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3,
                                                int j)
{
  const uint32_t lo = __builtin_amdgcn_perm(d1, d0, 0x0C0C0000u | ((4u + j) << 8) | (uint32_t)j);
  const uint32_t hi = __builtin_amdgcn_perm(d3, d2, ((4u + j) << 24) | ((uint32_t)j << 16) | 0x0C0Cu);
  return lo | hi;
}

template <int BITS, int G = kPassGroup>
__device__ __forceinline__ void pass_m16_packed(int16_t *P, int g, const uint32_t *v, long alo,
                                                long ahi, v4i bias)
{
  static_assert(G == 4, "four columns per group");
  constexpr uint32_t kFieldMask = ((1u << BITS) - 1u) * 0x01010101u;
  const v4i zero = {0, 0, 0, 0};
  int16_t *rows = P + 64 * g;  // row 4g; row 4g + r at +16r
#pragma unroll
  for (int cq = 0; cq < 4; cq++) {
    long b[4];
#pragma unroll
    for (int h = 0; h < 4; h++) {
      uint32_t gb;
      int sh;
      if (BITS == 2) {  // row r = dword r; column c at byte c/4, bit 2(c%4)
        gb = gather_byte(v[0], v[1], v[2], v[3], cq);
        sh = 2 * h;
      } else {  // row r = dwords 2r, 2r+1; column c at row byte c/2, bit BITS*(c%2)
        const int rb = 2 * cq + (h >> 1), dw = rb >> 2, by = rb & 3;
        gb = gather_byte(v[dw], v[2 + dw], v[4 + dw], v[6 + dw], by);
        sh = BITS * (h & 1);
      }
      const uint32_t f = (gb >> sh) & kFieldMask;
      const uint32_t sgn = (gb >> (sh + BITS - 1)) & 0x01010101u;
      b[h] = (long)(((uint64_t)sgn << 32) | f);
    }
    v4i lo[4], hi[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      lo[u] = mfma16(alo, b[u], bias);
#pragma unroll
    for (int u = 0; u < 4; u++)
      hi[u] = mfma16(ahi, b[u], zero);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      uint32_t val[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        val[u] = reduce_2q(lo[u][r], hi[u][r]);
      st2(rows + 16 * r + 4 * cq, pack2(val[0], val[1]), pack2(val[2], val[3]));
    }
    __builtin_amdgcn_sched_barrier(0);  // one column quad's operands live at a time
  }
}

// Contraction over c inside each block m (one matrix per block), G blocks
// per group.
template <int MODE, class ALo, class AHi, int G = kPassGroup>
__device__ __forceinline__ void pass_c16(int16_t *P, int g, ALo alo, AHi ahi,
                                         const int32_t (*s_bias)[16])
{
  const v4i zero = {0, 0, 0, 0};
#pragma unroll
  for (int m0 = 0; m0 < 16; m0 += G) {
    long b[G];
    v4i bias[G], lo[G], hi[G];
#pragma unroll
    for (int u = 0; u < G; u++) {
      const uint2 d = ld2(P + 16 * (m0 + u) + 4 * g);
      b[u] = digits8(d.x, d.y);
      bias[u] = *reinterpret_cast<const v4i *>(&s_bias[m0 + u][4 * g]);
    }
#pragma unroll
    for (int u = 0; u < G; u++)
      lo[u] = mfma16(alo(m0 + u), b[u], bias[u]);
#pragma unroll
    for (int u = 0; u < G; u++)
      hi[u] = mfma16(ahi(m0 + u), b[u], zero);
#pragma unroll
    for (int u = 0; u < G; u++) {
      uint32_t r[4];
#pragma unroll
      for (int i = 0; i < 4; i++)
        r[i] = reduce_2q(lo[u][i], hi[u][i]);
      st2(P + 16 * (m0 + u) + 4 * g, finish2<MODE>(pack2(r[0], r[1])),
          finish2<MODE>(pack2(r[2], r[3])));
    }
    // One group's operands live at a time: left free, the scheduler may
    // hoist every group's LDS reads and spill.
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ---------------------------------------------------------------------------
// 32x32x32 passes: lane l = 32h + t works on poly t; its K group and output
// rows are U_h = {4h..4h+3, 8+4h..8+4h+3}; accumulator register q (< 8) is
// the lo limb of row U_h[q], register q+8 its hi limb.
// ---------------------------------------------------------------------------
__device__ __forceinline__ v16i acc_init(const v4i &b0, const v4i &b1)
{
  return v16i{b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3], 0, 0, 0, 0, 0, 0, 0, 0};
}

__device__ __forceinline__ v4i op32(long a, long b)
{
  // [x0' rows U[0..3] | x0' rows U[4..7] | x1' rows U[0..3] | x1' rows U[4..7]]
  return v4i{(int)(uint32_t)a, (int)(uint32_t)b, (int)(uint32_t)((uint64_t)a >> 32),
             (int)(uint32_t)((uint64_t)b >> 32)};
}

template <int MODE>
__device__ __forceinline__ void pass_m32(int16_t *P, int h, v4i a, const v4i &b0, const v4i &b1)
{
  int16_t *ra = P + 64 * h;       // rows 4h .. 4h+3
  int16_t *rb = P + 128 + 64 * h; // rows 8+4h .. 8+4h+3
  const v16i acc0 = acc_init(b0, b1);
#pragma unroll
  for (int cq = 0; cq < 4; cq++) {
    uint2 RA[4], RB[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      RA[r] = ld2(ra + 16 * r + 4 * cq);
      RB[r] = ld2(rb + 16 * r + 4 * cq);
    }
    long ba[4], bb[4];
    col_digits4(RA, ba);
    col_digits4(RB, bb);
    v16i acc[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      acc[u] = mfma32(a, op32(ba[u], bb[u]), acc0);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      uint32_t va[4], vb[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        va[u] = reduce_2q(acc[u][r], acc[u][r + 8]);
        vb[u] = reduce_2q(acc[u][r + 4], acc[u][r + 12]);
      }
      st2(ra + 16 * r + 4 * cq, finish2<MODE>(pack2(va[0], va[1])),
          finish2<MODE>(pack2(va[2], va[3])));
      st2(rb + 16 * r + 4 * cq, finish2<MODE>(pack2(vb[0], vb[1])),
          finish2<MODE>(pack2(vb[2], vb[3])));
    }
  }
}

template <int MODE, class A>
__device__ __forceinline__ void pass_c32(int16_t *P, int h, A a, const int32_t (*s_bias)[16])
{
#pragma unroll
  for (int m0 = 0; m0 < 16; m0 += 4) {
    v16i acc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int m = m0 + u;
      const uint2 x0 = ld2(P + 16 * m + 4 * h), x1 = ld2(P + 16 * m + 8 + 4 * h);
      const v4i c0 = *reinterpret_cast<const v4i *>(&s_bias[m][4 * h]);
      const v4i c1 = *reinterpret_cast<const v4i *>(&s_bias[m][8 + 4 * h]);
      acc[u] = mfma32(a(m), op32(digits8(x0.x, x0.y), digits8(x1.x, x1.y)), acc_init(c0, c1));
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int m = m0 + u;
      uint32_t r[8];
#pragma unroll
      for (int q = 0; q < 8; q++)
        r[q] = reduce_2q(acc[u][q], acc[u][q + 8]);
      st2(P + 16 * m + 4 * h, finish2<MODE>(pack2(r[0], r[1])), finish2<MODE>(pack2(r[2], r[3])));
      st2(P + 16 * m + 8 + 4 * h, finish2<MODE>(pack2(r[4], r[5])),
          finish2<MODE>(pack2(r[6], r[7])));
    }
  }
}

}  // namespace tk
