// Host constant builder: twiddles and the MFMA operand matrices, all from
// closed forms in plain modular arithmetic.
//
// The reference builds its matrices by pushing unit vectors through host
// copies of its Montgomery layer code (ntt_kernels.cuh:842-925).  Here the
// stage maps are obtained by pushing unit vectors through a mod-q model of the
// same butterfly network, where a Montgomery product fqmul(zeta, x) with
// zeta = 17^brv7(k) * 2^16 is the plain product 17^brv7(k) * x
// (ref/ntt.c:68-70, ref/reduce.c:16-23).  The unit tests check these maps
// against their closed forms (Vandermonde, parity-split blocks, inverse
// scales) and against the reference transform end to end.
#include <string.h>

#include "tk_internal.h"

namespace tk {
namespace {

unsigned brv(unsigned x, int bits)
{
  unsigned r = 0;
  for (int b = 0; b < bits; b++)
    r |= ((x >> b) & 1u) << (bits - 1 - b);
  return r;
}

int64_t modq(int64_t v)
{
  v %= kQ;
  return v < 0 ? v + kQ : v;
}

int64_t powq(int64_t b, unsigned e)
{
  int64_t r = 1;
  b = modq(b);
  while (e) {
    if (e & 1u)
      r = r * b % kQ;
    b = b * b % kQ;
    e >>= 1;
  }
  return r;
}

// omega_k = 17^brv7(k): the plain-domain twiddle of zetas[k].
int64_t omega(unsigned k) { return powq(17, brv(k, 7)); }

// Forward Cooley-Tukey layers from length len_hi down to len_lo; k0 is the
// index of the first twiddle those layers consume (ref/ntt.c:85-94).
void fwd_layers(int64_t r[256], unsigned len_hi, unsigned len_lo, unsigned k0)
{
  unsigned k = k0;
  for (unsigned len = len_hi; len >= len_lo; len >>= 1)
    for (unsigned start = 0; start < 256; start += 2 * len) {
      const int64_t w = omega(k++);
      for (unsigned j = start; j < start + len; j++) {
        const int64_t t = w * r[j + len] % kQ;
        r[j + len] = modq(r[j] - t);
        r[j] = modq(r[j] + t);
      }
    }
}

// Inverse Gentleman-Sande layers from len_lo up to len_hi, twiddles counting
// down from k0 (ref/ntt.c:112-122).
void inv_layers(int64_t r[256], unsigned len_lo, unsigned len_hi, unsigned k0)
{
  unsigned k = k0;
  for (unsigned len = len_lo; len <= len_hi; len <<= 1)
    for (unsigned start = 0; start < 256; start += 2 * len) {
      const int64_t w = omega(k--);
      for (unsigned j = start; j < start + len; j++) {
        const int64_t t = r[j];
        r[j] = modq(t + r[j + len]);
        r[j + len] = modq(w * (r[j + len] - t));
      }
    }
}

int centred(int64_t v)
{
  v = modq(v);
  return (int)(v > kQ / 2 ? v - kQ : v);
}

// Signed limbs of a centred entry: v = v0 + 256*v1, v0 in [-128,127].
void limbs(int v, int8_t *v0, int8_t *v1)
{
  const int lo = (int8_t)(uint8_t)(v & 255);
  *v0 = (int8_t)lo;
  *v1 = (int8_t)((v - lo) / 256);
}

// V[i][d][u]: d = 0 -> W[i][u], d = 1 -> w1*W[i][u] mod q; centred.  w1 is
// the weight of the second data digit: 256 for the byte digits of any int16,
// -2^b for the (field, sign bit) digits of packed b-bit coefficients.
void digit_matrix(const uint16_t W[16][16], int V[16][2][16], int64_t w1 = 256)
{
  for (int i = 0; i < 16; i++)
    for (int u = 0; u < 16; u++) {
      V[i][0][u] = centred(W[i][u]);
      V[i][1][u] = centred(w1 * (int64_t)W[i][u]);
    }
}

void pack16(const uint16_t W[16][16], uint32_t out[2][64][2], int64_t w1 = 256)
{
  int V[16][2][16];
  digit_matrix(W, V, w1);
  for (int lane = 0; lane < 64; lane++) {
    const int g = lane >> 4, i = lane & 15;
    uint8_t b0[8], b1[8];
    for (int b = 0; b < 8; b++) {
      int8_t l0, l1;
      limbs(V[i][b / 4][4 * g + b % 4], &l0, &l1);
      b0[b] = (uint8_t)l0;
      b1[b] = (uint8_t)l1;
    }
    memcpy(out[0][lane], b0, 8);
    memcpy(out[1][lane], b1, 8);
  }
}

int u_set32(int h, int q) { return q < 4 ? 4 * h + q : 8 + 4 * h + (q - 4); }

void pack32(const uint16_t W[16][16], uint32_t out[64][4])
{
  int V[16][2][16];
  digit_matrix(W, V);
  for (int lane = 0; lane < 64; lane++) {
    const int h = lane >> 5, rho = lane & 31, i = rho & 15, hi = rho >> 4;
    uint8_t by[16];
    for (int b = 0; b < 16; b++) {
      int8_t l0, l1;
      limbs(V[i][b / 8][u_set32(h, b % 8)], &l0, &l1);
      by[b] = (uint8_t)(hi ? l1 : l0);
    }
    memcpy(out[lane], by, 16);
  }
}

// offset: the constant the digits leave out of every data value (128 for
// the byte digits, 0 for packed digits), which contributes offset * sum_u W.
void row_bias(const uint16_t W[16][16], int extra, int32_t out[16], int offset = 128)
{
  for (int i = 0; i < 16; i++) {
    int64_t s = 0;
    for (int u = 0; u < 16; u++)
      s += W[i][u];
    out[i] = (int32_t)(modq(offset * s) + (int64_t)kBiasQ * kQ + extra);
  }
}

}  // namespace

void build_tables(HostMatrices *hm, DeviceTables *dt)
{
  memset(hm, 0, sizeof(*hm));
  memset(dt, 0, sizeof(*dt));

  // zetas[i] = centred(17^brv7(i) * 2^16 mod q): the reference table
  // (ref/ntt.c:39-56; generator documented at ref/ntt.c:6-37).
  for (unsigned i = 0; i < 128; i++) {
    const int z = centred(omega(i) * 65536);
    hm->zetas[i] = (int16_t)z;
    dt->zetas[i] = (int16_t)z;
    dt->zetas_qinv[i] = (int16_t)(uint16_t)((uint32_t)(z * kQinv) & 0xffffu);
  }

  int64_t r[256];
  // Layers 1-4 act on m = index >> 4 identically for every c: probe e_{16m}.
  for (int m = 0; m < 16; m++) {
    memset(r, 0, sizeof(r));
    r[16 * m] = 1;
    fwd_layers(r, 128, 16, 1);
    for (int i = 0; i < 16; i++)
      hm->fwd1[i][m] = (uint16_t)r[16 * i];
    memset(r, 0, sizeof(r));
    r[16 * m] = 1;
    inv_layers(r, 16, 128, 15);
    for (int i = 0; i < 16; i++)
      hm->inv1[i][m] = (uint16_t)modq(r[16 * i] * 512);  // fqmul(.,1441) == x * 2^9
  }
  // Layers 5-7 stay inside each 16-coefficient block: probe e_{16b+j}.
  for (int b = 0; b < 16; b++)
    for (int j = 0; j < 16; j++) {
      memset(r, 0, sizeof(r));
      r[16 * b + j] = 1;
      fwd_layers(r, 8, 2, 16);
      for (int c = 0; c < 16; c++)
        hm->fwd2[b][c][j] = (uint16_t)r[16 * b + c];
      memset(r, 0, sizeof(r));
      r[16 * b + j] = 1;
      inv_layers(r, 2, 8, 127);
      for (int c = 0; c < 16; c++)
        hm->inv2[b][c][j] = (uint16_t)r[16 * b + c];
    }

  // 2^-16 mod q = 169 (65536 * 169 = 3329 * 3327 + 1)
  for (int i = 0; i < 16; i++)
    for (int m = 0; m < 16; m++)
      hm->inv1p[i][m] = (uint16_t)modq((int64_t)hm->inv1[i][m] * 169);
  for (int i = 0; i < 64; i++)
    dt->gamma[i] = (uint16_t)omega(64 + i);

  pack16(hm->fwd1, dt->m16_fwd1[0]);
  pack16(hm->inv1p, dt->m16_inv1p[0]);
  row_bias(hm->inv1p, kCentre, dt->bias_inv1p[0]);
  pack16(hm->inv1, dt->m16_inv1[0]);
  pack32(hm->fwd1, dt->m32_fwd1[0]);
  pack32(hm->inv1, dt->m32_inv1[0]);
  row_bias(hm->fwd1, 0, dt->bias_fwd1[0]);
  // Packed b-bit input: digits (field f, sign bit s) with x = f - 2^b s.
  for (int b = 2; b <= 4; b++)
    pack16(hm->fwd1, dt->m16_fwd1pk[b - 2], -(int64_t)(1 << b));
  row_bias(hm->fwd1, 0, dt->bias_fwd1pk[0], 0);
  row_bias(hm->inv1, kCentre, dt->bias_inv1[0]);
  for (int b = 0; b < 16; b++) {
    pack16(hm->fwd2[b], dt->m16_fwd2[b]);
    pack16(hm->inv2[b], dt->m16_inv2[b]);
    pack32(hm->fwd2[b], dt->m32_fwd2[b]);
    pack32(hm->inv2[b], dt->m32_inv2[b]);
    row_bias(hm->fwd2[b], 0, dt->bias_fwd2[b]);
    row_bias(hm->inv2[b], 0, dt->bias_inv2[b]);
  }
}

}  // namespace tk
