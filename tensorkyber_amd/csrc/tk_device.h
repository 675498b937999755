// Device-side helpers shared by the gfx950 kernels: exact reference
// reductions, the MFMA digit/limb arithmetic, wave-level synchronisation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tk_internal.h"

namespace tk {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Streaming (nontemporal) 16-byte accesses for data that is touched once:
// polys are read once and written once per pass, so they bypass cache
// retention (about 8-10% more HBM bandwidth than plain accesses here).
__device__ __forceinline__ v4i ld_stream(const v4i *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(v4i *p, v4i v) { __builtin_nontemporal_store(v, p); }

// ---- exact reference arithmetic (int16 narrowing as in the C reference) ----

// ref/reduce.c:16-23 with the product's QINV factor pre-multiplied into the
// twiddle: (int16)((int16)(z*x) * QINV) == (int16)(x * zq), zq = z*QINV mod 2^16.
__device__ __forceinline__ int16_t fqmul_zq(int16_t x, int16_t z, int16_t zq)
{
  const int32_t a = (int32_t)z * x;
  const int16_t t = (int16_t)((int32_t)x * zq);
  return (int16_t)((a - (int32_t)t * kQ) >> 16);
}

__device__ __forceinline__ int16_t fqmul(int16_t a, int16_t b)
{
  const int32_t p = (int32_t)a * b;
  const int16_t t = (int16_t)((int32_t)(int16_t)p * kQinv);
  return (int16_t)((p - (int32_t)t * kQ) >> 16);
}

// ref/reduce.c:35-42
__device__ __forceinline__ int16_t barrett16(int16_t a)
{
  int16_t t = (int16_t)(((int32_t)kBarrettV * a + (1 << 25)) >> 26);
  t = (int16_t)(t * kQ);
  return (int16_t)(a - t);
}

// Canonical [0, q) and centred [-(q-1)/2, (q-1)/2] of any int16.
__device__ __forceinline__ int16_t canon16(int16_t a)
{
  const int16_t b = barrett16(a);  // centred
  return (int16_t)(b < 0 ? b + kQ : b);
}
__device__ __forceinline__ int16_t centre16(int16_t a) { return barrett16(a); }

// ---- MFMA digit arithmetic -------------------------------------------------

// v_mul_hi_u32_u24: bits 47..32 of the 48-bit product of the low 24 bits.
__device__ __forceinline__ uint32_t mulhi_u24(uint32_t a, uint32_t b)
{
  uint32_t r;
  asm("v_mul_hi_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// v_mad_i32_i24: a * b + c with a, b taken as signed 24-bit values.  Written
// out so the subtract of the Barrett step stays one full-rate instruction.
__device__ __forceinline__ uint32_t mad_i24(uint32_t a, int b, uint32_t c)
{
  uint32_t r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}

// Digits of four int16 values held as two packed pairs d0 = (x0,x1),
// d1 = (x2,x3): x = 128 + x0' + 256*x1' with x0' = (x & 255) - 128 (byte 0
// with its top bit flipped) and x1' = x >> 8 (byte 1).  Returns the 8-byte
// MFMA operand [x0'(4) | x1'(4)] as a 64-bit value.
__device__ __forceinline__ uint32_t digit_lo(uint32_t d0, uint32_t d1)
{
  return __builtin_amdgcn_perm(d1, d0, 0x06040200u) ^ 0x80808080u;
}
__device__ __forceinline__ uint32_t digit_hi(uint32_t d0, uint32_t d1)
{
  return __builtin_amdgcn_perm(d1, d0, 0x07050301u);
}
__device__ __forceinline__ long digits8(uint32_t d0, uint32_t d1)
{
  return (long)(((uint64_t)digit_hi(d0, d1) << 32) | digit_lo(d0, d1));
}

// Pack the low halves of two 32-bit values into one (a, b) int16 pair.
__device__ __forceinline__ uint32_t pack2(uint32_t a, uint32_t b)
{
  return __builtin_amdgcn_perm(b, a, 0x05040100u);
}

// Combine the two limb accumulators and reduce: S' = lo + 256*hi (the row
// bias, folded into lo by the MFMA's C input, keeps 0 <= S' < 2^24), then
// r = S' - q*floor~(S'/q) in [0, 2q).
__device__ __forceinline__ uint32_t reduce_2q(int lo, int hi)
{
  const uint32_t s = (uint32_t)lo + ((uint32_t)hi << 8);
  const uint32_t t = mulhi_u24(s, kBarrettM);
  return mad_i24(t, -kQ, s);
}

__device__ __forceinline__ uint32_t reduce_q(int lo, int hi)
{
  const uint32_t r = reduce_2q(lo, hi);
  return min(r, r - (uint32_t)kQ);  // [0, q)
}

__device__ __forceinline__ v4i mfma16(long a, long b, v4i c)
{
  return __builtin_amdgcn_mfma_i32_16x16x32_i8(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v16i mfma32(v4i a, v4i b, v16i c)
{
  return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// Wait until every vector memory operation the wave has issued is done
// (s_waitcnt vmcnt(0), expcnt and lgkmcnt left at their maximum).
//
// Kernels call it once, after the loads that precede their tile loop.  The
// compiler places one wait per use of a loaded register, valid on every path
// into it: at the loop header that covers both the entry edge, where a
// prologue load may be the last operation issued, and the back edge, where
// the previous tile's stores follow the loads.  Vector memory operations
// complete in issue order, so a wait sized for the entry edge (vmcnt(0))
// also waits for those stores on every iteration.  Waiting here, once, keeps
// the loop's waits sized for the back edge.
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// All lanes of the wave have finished their LDS writes before any lane reads
// what another lane wrote.  One wave owns each LDS region, so no workgroup
// barrier is needed.
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace tk
