/* Self-check of the oracle under AddressSanitizer and UndefinedBehaviorSanitizer
 * (test infrastructure: built by `make -C oracle sanitize-check` with gcc,
 * host code only, and run by tests/test_oracle.py).
 *
 * Every exported function runs on small batches that span the int16 range
 * (the reference relies on implementation-defined int16 narrowing, which is
 * not undefined behaviour; any undefined operation or out-of-bounds access
 * aborts the run), and a few invariants are checked on the way:
 *   - the exhaustive reduction checks of oracle_check_reductions;
 *   - invntt(ntt(x)) == 2^16 x (mod q) for |x| <= q;
 *   - basemul_batch equals poly_basemul per poly;
 *   - basemul(a, b) == a (mod q) for b = (2^16 mod q, 0) on every pair;
 *   - pack/unpack round trips for 2, 3 and 4 bits;
 *   - keygen -> encrypt -> decrypt recovers every message bit.
 * Prints "sanitize-check ok" and exits 0 on success. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kyber_oracle.h"

#define N ORACLE_N
#define Q ORACLE_Q

static uint64_t g_state = 0x9e3779b97f4a7c15ull;

static uint64_t next64(void)
{
  /* SplitMix64 */
  uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static int16_t draw(int lo, int hi) /* in [lo, hi) */
{
  return (int16_t)(lo + (int)(next64() % (uint64_t)(hi - lo)));
}

static void fill(int16_t *a, size_t n, int lo, int hi)
{
  for (size_t i = 0; i < n; i++)
    a[i] = draw(lo, hi);
}

static int canon(int v)
{
  int r = v % Q;
  return r < 0 ? r + Q : r;
}

static int fail(const char *what)
{
  fprintf(stderr, "sanitize-check: %s\n", what);
  return 1;
}

int main(void)
{
  enum { NP = 24, K = 3, NI = 5 };
  if (oracle_check_reductions() != 0)
    return fail("reductions");

  /* ntt / invntt over the full int16 range (exercises the wrapping paths),
   * and the round trip on |x| <= q. */
  int16_t *x = malloc(sizeof(int16_t) * NP * N), *y = malloc(sizeof(int16_t) * NP * N);
  fill(x, (size_t)NP * N, -32768, 32768);
  memcpy(y, x, sizeof(int16_t) * NP * N);
  oracle_ntt_batch(y, NP, 1);
  oracle_invntt_batch(y, NP, 1);
  fill(x, (size_t)NP * N, -Q, Q + 1);
  memcpy(y, x, sizeof(int16_t) * NP * N);
  oracle_roundtrip_batch(y, NP, 1);
  for (size_t i = 0; i < (size_t)NP * N; i++)
    if (canon(y[i]) != canon(x[i] * 65536 % Q))
      return fail("round trip");

  /* basemul */
  int16_t *a = malloc(sizeof(int16_t) * NP * N), *b = malloc(sizeof(int16_t) * NP * N);
  int16_t *r = malloc(sizeof(int16_t) * NP * N), r1[N];
  fill(a, (size_t)NP * N, -32768, 32768);
  fill(b, (size_t)NP * N, -32768, 32768);
  oracle_basemul_batch(r, a, b, NP, 1);
  for (int p = 0; p < NP; p++) {
    oracle_poly_basemul(r1, a + p * N, b + p * N);
    if (memcmp(r1, r + p * N, sizeof r1))
      return fail("basemul batch");
  }
  for (int i = 0; i < NP * N; i++)
    b[i] = (int16_t)(i % 2 == 0 ? 65536 % Q : 0);
  oracle_basemul_batch(r, a, b, NP, 1);
  for (int i = 0; i < NP * N; i++)
    if (canon(r[i]) != canon(a[i]))
      return fail("basemul identity");

  /* packing */
  for (int bits = 2; bits <= 4; bits++) {
    const size_t pb = oracle_packed_poly_bytes(bits);
    uint8_t *pk = malloc(pb * NP);
    const int lo = -(1 << (bits - 1)), hi = 1 << (bits - 1);
    fill(x, (size_t)NP * N, lo, hi);
    oracle_pack_small(pk, x, NP, bits);
    oracle_unpack_small(y, pk, NP, bits);
    if (memcmp(x, y, sizeof(int16_t) * NP * N))
      return fail("pack round trip");
    free(pk);
  }

  /* matvec and the K-PKE cores; decrypt recovers the message */
  const size_t vec = (size_t)NI * K * N, mat = vec * K, one = (size_t)NI * N;
  int16_t *A = malloc(sizeof(int16_t) * mat), *At = malloc(sizeof(int16_t) * mat);
  int16_t *s = malloc(sizeof(int16_t) * vec), *e = malloc(sizeof(int16_t) * vec);
  int16_t *t = malloc(sizeof(int16_t) * vec), *th = malloc(sizeof(int16_t) * vec);
  int16_t *sh = malloc(sizeof(int16_t) * vec), *rr = malloc(sizeof(int16_t) * vec);
  int16_t *e1 = malloc(sizeof(int16_t) * vec), *u = malloc(sizeof(int16_t) * vec);
  int16_t *e2 = malloc(sizeof(int16_t) * one), *m = malloc(sizeof(int16_t) * one);
  int16_t *v = malloc(sizeof(int16_t) * one), *w = malloc(sizeof(int16_t) * one);
  fill(A, mat, 0, Q);
  for (int in = 0; in < NI; in++)
    for (int i = 0; i < K; i++)
      for (int j = 0; j < K; j++)
        memcpy(At + (((size_t)in * K + i) * K + j) * N, A + (((size_t)in * K + j) * K + i) * N,
               sizeof(int16_t) * N);
  fill(s, vec, -2, 3);
  fill(e, vec, -2, 3);
  fill(rr, vec, -2, 3);
  fill(e1, vec, -2, 3);
  fill(e2, one, -2, 3);
  for (size_t i = 0; i < one; i++)
    m[i] = (int16_t)((next64() & 1) ? (Q + 1) / 2 : 0);
  oracle_matvec_batch(t, A, s, K, NI, 1);
  oracle_kpke_keygen_batch(th, sh, A, s, e, K, NI, 1);
  oracle_kpke_encrypt_batch(u, v, At, th, rr, e1, e2, m, K, NI, 1);
  oracle_kpke_decrypt_batch(w, sh, u, v, K, NI, 1);
  for (size_t i = 0; i < one; i++) {
    const int d = canon(w[i]);
    const int bit = d > Q / 4 && d < 3 * Q / 4;
    if (bit != (m[i] != 0))
      return fail("message round trip");
  }

  free(x); free(y); free(a); free(b); free(r);
  free(A); free(At); free(s); free(e); free(t); free(th); free(sh); free(rr); free(e1);
  free(u); free(e2); free(m); free(v); free(w);
  printf("sanitize-check ok\n");
  return 0;
}
