// Internal declarations shared by the host constant builder, the kernels and
// the C ABI.  Not installed; the public surface is include/tkyber.h.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace tk {

constexpr int kN = 256;
constexpr int kQ = 3329;
constexpr int kQinv = -3327;       // q^-1 mod 2^16 (ref/reduce.h:8)
constexpr int kMont = -1044;       // 2^16 mod q    (ref/reduce.h:7)
constexpr int kBarrettV = 20159;   // round(2^26 / q) (ref/reduce.c:37)
constexpr int kInvScale = 1441;    // 2^32 / 128 mod q (ref/ntt.c:109)

// Unsigned 24-bit Barrett used by every MFMA epilogue: for 0 <= S < 2^24,
// t = mulhi_u24(S, kBarrettM) is floor(S/q) or one less, so S - t*q is in
// [0, 2q).  kBarrettM = floor(2^32 / q).
constexpr uint32_t kBarrettM = 1290167u;
// Positive offset (a multiple of q) folded into every MFMA row bias so the
// combined accumulator S' = C_lo + 256*C_hi + bias is never negative:
// |C_lo + 256*C_hi| <= 32*128*128 + 256*32*7*128 = 7,864,320 < kBiasQ*q, and
// S' <= 7,864,320 + q + 1664 + kBiasQ*q < 2^24 = 16,777,216.
constexpr int kBiasQ = 2500;
constexpr int kCentre = (kQ - 1) / 2;  // 1664

// Output forms of the kernels' last pass (kFormInter: intermediate [0, 2q)).
enum OutForm { kFormInter = 0, kFormCanon = 1, kFormCentred = 2, kFormRef = 3 };

// Stage indices of the four constant matrices.
enum Stage { kFwd1 = 0, kFwd2 = 1, kInv2 = 2, kInv1 = 3 };

// Per-lane MFMA A-operand tables, uploaded once per device.  For every
// constant matrix W (16x16, entries mod q) the MFMA contracts
//   V = [ W | 256*W mod q ]  (16 x 32, centred entries, |V| <= 1664)
// against the data digits [x0' ; x1'] where x = 128 + x0' + 256*x1',
// x0' = (x & 255) - 128, x1' = x >> 8 (exact for every int16 x).  Each V is
// split into signed limbs V = v0 + 256*v1, v0 in [-128,127], v1 in [-7,7];
// the offset 128 contributes the per-row constant 128*sum_u W[i][u], kept in
// the row bias together with kBiasQ*q (and kCentre for centred outputs).
//
// MFMA16 (v_mfma_i32_16x16x32_i8): lane l = 16g + i holds row i; byte b of
// its 8-byte operand is K slot (digit b/4, source index u = 4g + b%4).
// MFMA32 (v_mfma_i32_32x32x32_i8): lane l = 32h + rho holds row rho (rows
// 0-15 carry v0, rows 16-31 carry v1 of row rho-16); byte b of its 16-byte
// operand is K slot (digit b/8, u = U_h[b%8]) with U_h = {4h..4h+3,
// 8+4h..8+4h+3}, the same index set that lane half h receives as output rows.
struct DeviceTables {
  int16_t zetas[128];       // Montgomery form, centred (the reference table)
  int16_t zetas_qinv[128];  // zetas[i] * QINV mod 2^16, sign-extended
  // MFMA16 operands: [matrix][limb][lane][2 dwords]
  uint32_t m16_fwd1[1][2][64][2];
  uint32_t m16_fwd2[16][2][64][2];
  uint32_t m16_inv2[16][2][64][2];
  uint32_t m16_inv1[1][2][64][2];
  // Inverse layers 4-1 without the Montgomery factor: plain NTT^-1 =
  // inv1 * 2^-16, used where the input carries no 2^-16 (fused products).
  uint32_t m16_inv1p[1][2][64][2];
  // Forward layers 1-4 for packed b-bit input (b = 2, 3, 4 at [b-2]): the
  // digits are the unsigned field f and its sign bit s, x = f - 2^b s, so
  // V = [ W | -2^b W mod q ]; the row bias carries no 128 offset.
  uint32_t m16_fwd1pk[3][2][64][2];
  // MFMA32 operands: [matrix][lane][4 dwords]
  uint32_t m32_fwd1[1][64][4];
  uint32_t m32_fwd2[16][64][4];
  uint32_t m32_inv2[16][64][4];
  uint32_t m32_inv1[1][64][4];
  // Row biases [matrix][row]; final stages include their output offset.
  int32_t bias_fwd1[1][16];
  int32_t bias_fwd2[16][16];
  int32_t bias_inv2[16][16];
  int32_t bias_inv1[1][16];
  int32_t bias_inv1p[1][16];  // includes kCentre
  int32_t bias_fwd1pk[1][16];  // kBiasQ * q only
  // Plain-domain basemul twiddles gamma_i = zetas[64+i] * 2^-16 mod q in [0, q)
  // (pair (4i,4i+1) uses gamma_i, pair (4i+2,4i+3) uses q - gamma_i).
  uint16_t gamma[64];
};

// Host-side tables (plain mod-q matrices) kept for introspection.
struct HostMatrices {
  int16_t zetas[128];
  uint16_t fwd1[16][16];      // [row i][m]   : layers 1-4
  uint16_t fwd2[16][16][16];  // [block][c][j]: layers 5-7
  uint16_t inv2[16][16][16];  // [block][c][j]: inverse layers len 2,4,8
  uint16_t inv1[16][16];      // [row i][m]   : inverse layers len 16..128, x1441
  uint16_t inv1p[16][16];     // inv1 * 2^-16 mod q: plain inverse of layers 1-4
};

// Builds both table sets from closed forms (tk_tables.cpp).  Pure host code.
void build_tables(HostMatrices *hm, DeviceTables *dt);

}  // namespace tk
