/*
 * kyber_oracle.c -- CPU restatement of the reference ML-KEM NTT path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product); see
 * kyber_oracle.h for the reference file:line each function follows and for
 * the checks that pin it.
 *
 * Narrowing conversions to int16_t are written as explicit casts where the
 * reference relies on the implicit narrowing of an assignment, a parameter or
 * a return value; gcc defines both as reduction modulo 2^16.
 */
#include "kyber_oracle.h"

#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define QINV (-3327) /* q^-1 mod 2^16, ref/reduce.h:8 */
#define MONT (-1044) /* 2^16 mod q,    ref/reduce.h:7 */

/* ref/reduce.c:16-23: t = (int16)a * QINV narrowed to int16; (a - t*q) >> 16. */
int16_t oracle_montgomery_reduce(int32_t a)
{
  int16_t t = (int16_t)((int32_t)(int16_t)a * QINV);
  return (int16_t)((a - (int32_t)t * ORACLE_Q) >> 16);
}

/* ref/reduce.c:35-42: v = round(2^26/q) = 20159, t = (v*a + 2^25) >> 26, a - t*q
 * with t and t*q narrowed to int16 as in the reference. */
int16_t oracle_barrett_reduce(int16_t a)
{
  const int16_t v = (int16_t)(((1 << 26) + ORACLE_Q / 2) / ORACLE_Q);
  int16_t t = (int16_t)(((int32_t)v * a + (1 << 25)) >> 26);
  t = (int16_t)(t * ORACLE_Q);
  return (int16_t)(a - t);
}

/* ref/ntt.c:68-70 */
int16_t oracle_fqmul(int16_t a, int16_t b)
{
  return oracle_montgomery_reduce((int32_t)a * b);
}

/* ref/ntt.c:39-56 holds the table; ntt.c:6-37 documents its generator.  Here it
 * is built from the closed form zetas[i] = centred(17^brv7(i) * 2^16 mod q) with
 * plain modular arithmetic, independent of the Montgomery helpers above. */
static int16_t g_zetas[128];
static int g_zetas_ready;

static unsigned brv7(unsigned i)
{
  unsigned r = 0;
  for (int b = 0; b < 7; b++)
    r |= ((i >> b) & 1u) << (6 - b);
  return r;
}

static unsigned powmod(unsigned base, unsigned e)
{
  unsigned long long r = 1, b = base % ORACLE_Q;
  while (e) {
    if (e & 1u)
      r = r * b % ORACLE_Q;
    b = b * b % ORACLE_Q;
    e >>= 1;
  }
  return (unsigned)r;
}

static void zetas_init(void)
{
  if (g_zetas_ready)
    return;
  for (unsigned i = 0; i < 128; i++) {
    unsigned v = (unsigned)((unsigned long long)powmod(17, brv7(i)) * 65536u % ORACLE_Q);
    int c = (int)v;
    if (c > ORACLE_Q / 2)
      c -= ORACLE_Q;
    g_zetas[i] = (int16_t)c;
  }
  g_zetas_ready = 1;
}

const int16_t *oracle_zetas(void)
{
  zetas_init();
  return g_zetas;
}

/* ref/ntt.c:80-95: seven Cooley-Tukey layers, len 128 -> 2, zetas[1..127] in
 * order, no reduction of the outputs. */
void oracle_ntt(int16_t r[256])
{
  const int16_t *z = oracle_zetas();
  unsigned k = 1;
  for (unsigned len = 128; len >= 2; len >>= 1) {
    for (unsigned start = 0; start < 256; start += 2 * len) {
      const int16_t zeta = z[k++];
      for (unsigned j = start; j < start + len; j++) {
        const int16_t t = oracle_fqmul(zeta, r[j + len]);
        r[j + len] = (int16_t)(r[j] - t);
        r[j] = (int16_t)(r[j] + t);
      }
    }
  }
}

/* ref/ntt.c:106-126: seven Gentleman-Sande layers, len 2 -> 128, zetas[127..1]
 * in order, Barrett reduction of every sum, then fqmul by 1441 = 2^32/128 mod q. */
void oracle_invntt(int16_t r[256])
{
  const int16_t *z = oracle_zetas();
  const int16_t f = 1441;
  unsigned k = 127;
  for (unsigned len = 2; len <= 128; len <<= 1) {
    for (unsigned start = 0; start < 256; start += 2 * len) {
      const int16_t zeta = z[k--];
      for (unsigned j = start; j < start + len; j++) {
        const int16_t t = r[j];
        r[j] = oracle_barrett_reduce((int16_t)(t + r[j + len]));
        r[j + len] = (int16_t)(r[j + len] - t);
        r[j + len] = oracle_fqmul(zeta, r[j + len]);
      }
    }
  }
  for (unsigned j = 0; j < 256; j++)
    r[j] = oracle_fqmul(r[j], f);
}

/* ref/ntt.c:139-146 (the += are int16 narrowing additions). */
void oracle_basemul(int16_t r[2], const int16_t a[2], const int16_t b[2], int16_t zeta)
{
  int16_t r0 = oracle_fqmul(a[1], b[1]);
  r0 = oracle_fqmul(r0, zeta);
  r0 = (int16_t)(r0 + oracle_fqmul(a[0], b[0]));
  int16_t r1 = oracle_fqmul(a[0], b[1]);
  r1 = (int16_t)(r1 + oracle_fqmul(a[1], b[0]));
  r[0] = r0;
  r[1] = r1;
}

void oracle_poly_basemul(int16_t r[256], const int16_t a[256], const int16_t b[256])
{
  const int16_t *z = oracle_zetas();
  for (unsigned i = 0; i < 64; i++) {
    oracle_basemul(&r[4 * i], &a[4 * i], &b[4 * i], z[64 + i]);
    oracle_basemul(&r[4 * i + 2], &a[4 * i + 2], &b[4 * i + 2], (int16_t)-z[64 + i]);
  }
}

static int resolve_threads(int nthreads)
{
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

int oracle_max_threads(void)
{
  return resolve_threads(0);
}

void oracle_ntt_batch(int16_t *polys, size_t npoly, int nthreads)
{
  oracle_zetas();
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long p = 0; p < n; p++)
    oracle_ntt(polys + (size_t)p * ORACLE_N);
}

void oracle_invntt_batch(int16_t *polys, size_t npoly, int nthreads)
{
  oracle_zetas();
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long p = 0; p < n; p++)
    oracle_invntt(polys + (size_t)p * ORACLE_N);
}

void oracle_roundtrip_batch(int16_t *polys, size_t npoly, int nthreads)
{
  oracle_zetas();
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long p = 0; p < n; p++) {
    oracle_ntt(polys + (size_t)p * ORACLE_N);
    oracle_invntt(polys + (size_t)p * ORACLE_N);
  }
}

void oracle_basemul_batch(int16_t *r, const int16_t *a, const int16_t *b, size_t npoly,
                          int nthreads)
{
  oracle_zetas();
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long p = 0; p < n; p++)
    oracle_poly_basemul(r + (size_t)p * ORACLE_N, a + (size_t)p * ORACLE_N,
                        b + (size_t)p * ORACLE_N);
}

void oracle_matvec_batch(int16_t *t, const int16_t *A_hat, const int16_t *s, int k,
                         size_t ninst, int nthreads)
{
  oracle_zetas();
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t shat[4][ORACLE_N];
    int16_t tmp[ORACLE_N];
    const int16_t *A = A_hat + (size_t)inst * k * k * ORACLE_N;
    for (int j = 0; j < k; j++) {
      memcpy(shat[j], s + ((size_t)inst * k + j) * ORACLE_N, sizeof(shat[j]));
      oracle_ntt(shat[j]);
    }
    for (int i = 0; i < k; i++) {
      int16_t *acc = t + ((size_t)inst * k + i) * ORACLE_N;
      oracle_poly_basemul(acc, A + ((size_t)i * k + 0) * ORACLE_N, shat[0]);
      for (int j = 1; j < k; j++) {
        oracle_poly_basemul(tmp, A + ((size_t)i * k + j) * ORACLE_N, shat[j]);
        for (int c = 0; c < ORACLE_N; c++)
          acc[c] = (int16_t)(acc[c] + tmp[c]);
      }
      for (int c = 0; c < ORACLE_N; c++)
        acc[c] = oracle_barrett_reduce(acc[c]);
      oracle_invntt(acc);
    }
  }
}

/* Invariant checks of the reductions against what the mathematics fixes,
 * over their whole domains (returns the number of violations):
 *   montgomery: every a in [-q*2^15, q*2^15 - 1] -> r == a * 2^-16 (mod q),
 *               |r| < q (ref/reduce.c:5-15 precondition and contract);
 *   barrett:    every int16 a -> r == a (mod q), |r| <= (q-1)/2
 *               (ref/reduce.c:25-34). */
long long oracle_check_reductions(void)
{
  long long bad = 0;
  const int64_t rinv = 169; /* 2^-16 mod q: 2285 * 169 = 386165 = 116*3329 + 1 */
  const int32_t lo = -ORACLE_Q * 32768, hi = ORACLE_Q * 32768 - 1;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int64_t a = lo; a <= hi; a++) {
    const int16_t r = oracle_montgomery_reduce((int32_t)a);
    int64_t want = (a % ORACLE_Q) * rinv % ORACLE_Q;
    if (want < 0)
      want += ORACLE_Q;
    int64_t got = r % ORACLE_Q;
    if (got < 0)
      got += ORACLE_Q;
    if (got != want || r <= -ORACLE_Q || r >= ORACLE_Q)
      bad++;
  }
  for (int a = -32768; a <= 32767; a++) {
    const int16_t r = oracle_barrett_reduce((int16_t)a);
    if (((a - r) % ORACLE_Q) != 0 || r < -(ORACLE_Q - 1) / 2 || r > (ORACLE_Q - 1) / 2)
      bad++;
  }
  return bad;
}

/* ------------------------------------------------------------------------ */
/* Packed small-coefficient input (ntt_bench/ntt_bench.cu:749-757 packs on the
 * host; the device unpack at :325-348 sign-extends each field). */

size_t oracle_packed_poly_bytes(int bits)
{
  if (bits < 2 || bits > 4)
    return 0;
  return (size_t)ORACLE_N / (size_t)(8 / bits);
}

void oracle_pack_small(uint8_t *packed, const int16_t *polys, size_t npoly, int bits)
{
  const size_t ppb = oracle_packed_poly_bytes(bits);
  const int per = 8 / bits;
  if (ppb == 0)
    return;
  for (size_t p = 0; p < npoly; p++) {
    uint8_t *dst = packed + p * ppb;
    const int16_t *src = polys + p * ORACLE_N;
    for (size_t b = 0; b < ppb; b++)
      dst[b] = 0;
    for (int i = 0; i < ORACLE_N; i++)
      dst[i / per] |= (uint8_t)((src[i] & ((1 << bits) - 1)) << ((i % per) * bits));
  }
}

void oracle_unpack_small(int16_t *polys, const uint8_t *packed, size_t npoly, int bits)
{
  const size_t ppb = oracle_packed_poly_bytes(bits);
  const int per = 8 / bits;
  if (ppb == 0)
    return;
  for (size_t p = 0; p < npoly; p++) {
    const uint8_t *src = packed + p * ppb;
    int16_t *dst = polys + p * ORACLE_N;
    for (int i = 0; i < ORACLE_N; i++) {
      const int code = (src[i / per] >> ((i % per) * bits)) & ((1 << bits) - 1);
      /* sign-extend the bits-wide field */
      dst[i] = (int16_t)(code >= (1 << (bits - 1)) ? code - (1 << bits) : code);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* K-PKE arithmetic cores (pqcrystals-kyber ref/poly.c, ref/polyvec.c and
 * ref/indcpa.c, upstream of ref/ and not vendored; restated from the
 * published reference: poly_tomont's factor is 2^32 mod q = 1353). */

static void o_poly_reduce(int16_t r[ORACLE_N])
{
  for (int i = 0; i < ORACLE_N; i++)
    r[i] = oracle_barrett_reduce(r[i]);
}

static void o_poly_tomont(int16_t r[ORACLE_N])
{
  const int16_t f = (int16_t)((1ULL << 32) % ORACLE_Q);
  for (int i = 0; i < ORACLE_N; i++)
    r[i] = oracle_montgomery_reduce((int32_t)r[i] * f);
}

static void o_poly_ntt(int16_t r[ORACLE_N])
{
  oracle_ntt(r);
  o_poly_reduce(r);
}

/* polyvec_basemul_acc_montgomery: a and b are k consecutive polys. */
static void o_basemul_acc(int16_t r[ORACLE_N], const int16_t *a, const int16_t *b, int k)
{
  int16_t t[ORACLE_N];
  oracle_poly_basemul(r, a, b);
  for (int j = 1; j < k; j++) {
    oracle_poly_basemul(t, a + (size_t)j * ORACLE_N, b + (size_t)j * ORACLE_N);
    for (int c = 0; c < ORACLE_N; c++)
      r[c] = (int16_t)(r[c] + t[c]);
  }
  o_poly_reduce(r);
}

void oracle_kpke_keygen_batch(int16_t *t_hat, int16_t *s_hat, const int16_t *A_hat,
                              const int16_t *s, const int16_t *e, int k, size_t ninst,
                              int nthreads)
{
  oracle_zetas();
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t eh[4][ORACLE_N];
    int16_t *sh = s_hat + (size_t)inst * k * ORACLE_N;
    memcpy(sh, s + (size_t)inst * k * ORACLE_N, (size_t)k * ORACLE_N * sizeof(int16_t));
    for (int j = 0; j < k; j++) {
      o_poly_ntt(sh + (size_t)j * ORACLE_N);
      memcpy(eh[j], e + ((size_t)inst * k + j) * ORACLE_N, sizeof(eh[j]));
      o_poly_ntt(eh[j]);
    }
    for (int i = 0; i < k; i++) {
      int16_t *pk = t_hat + ((size_t)inst * k + i) * ORACLE_N;
      o_basemul_acc(pk, A_hat + ((size_t)inst * k + i) * k * ORACLE_N, sh, k);
      o_poly_tomont(pk);
      for (int c = 0; c < ORACLE_N; c++)
        pk[c] = (int16_t)(pk[c] + eh[i][c]);
      o_poly_reduce(pk);
    }
  }
}

void oracle_kpke_encrypt_batch(int16_t *u, int16_t *v, const int16_t *At_hat,
                               const int16_t *t_hat, const int16_t *r, const int16_t *e1,
                               const int16_t *e2, const int16_t *m, int k, size_t ninst,
                               int nthreads)
{
  oracle_zetas();
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t rh[4 * ORACLE_N];
    memcpy(rh, r + (size_t)inst * k * ORACLE_N, (size_t)k * ORACLE_N * sizeof(int16_t));
    for (int j = 0; j < k; j++)
      o_poly_ntt(rh + (size_t)j * ORACLE_N);
    for (int i = 0; i < k; i++) {
      int16_t *b = u + ((size_t)inst * k + i) * ORACLE_N;
      const int16_t *ei = e1 + ((size_t)inst * k + i) * ORACLE_N;
      o_basemul_acc(b, At_hat + ((size_t)inst * k + i) * k * ORACLE_N, rh, k);
      oracle_invntt(b);
      for (int c = 0; c < ORACLE_N; c++)
        b[c] = (int16_t)(b[c] + ei[c]);
      o_poly_reduce(b);
    }
    int16_t *vv = v + (size_t)inst * ORACLE_N;
    const int16_t *e2i = e2 + (size_t)inst * ORACLE_N;
    o_basemul_acc(vv, t_hat + (size_t)inst * k * ORACLE_N, rh, k);
    oracle_invntt(vv);
    for (int c = 0; c < ORACLE_N; c++)
      vv[c] = (int16_t)(vv[c] + e2i[c]);
    if (m) {
      const int16_t *mi = m + (size_t)inst * ORACLE_N;
      for (int c = 0; c < ORACLE_N; c++)
        vv[c] = (int16_t)(vv[c] + mi[c]);
    }
    o_poly_reduce(vv);
  }
}

void oracle_kpke_decrypt_batch(int16_t *w, const int16_t *s_hat, const int16_t *u,
                               const int16_t *v, int k, size_t ninst, int nthreads)
{
  oracle_zetas();
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(resolve_threads(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t bh[4 * ORACLE_N];
    memcpy(bh, u + (size_t)inst * k * ORACLE_N, (size_t)k * ORACLE_N * sizeof(int16_t));
    for (int j = 0; j < k; j++)
      o_poly_ntt(bh + (size_t)j * ORACLE_N);
    int16_t *mp = w + (size_t)inst * ORACLE_N;
    const int16_t *vi = v + (size_t)inst * ORACLE_N;
    o_basemul_acc(mp, s_hat + (size_t)inst * k * ORACLE_N, bh, k);
    oracle_invntt(mp);
    for (int c = 0; c < ORACLE_N; c++)
      mp[c] = (int16_t)(vi[c] - mp[c]);
    o_poly_reduce(mp);
  }
}
