/*
 * ref_driver.c -- batch entry points around the reference C sources.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
 * ntt_bench/ref/ntt.c and ntt_bench/ref/reduce.c, read where they lie in the
 * reference checkout, into oracle/_ref/libkyber_ref.so.  Nothing of the
 * reference is copied: this file only includes its header and loops over its
 * functions (ntt.h:7-17), so the golden-fixture script, the oracle tests and
 * the CPU baseline can call the real reference on whole batches.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ntt.h"
#include "reduce.h"

static int nthr(int n)
{
#ifdef _OPENMP
  return n > 0 ? n : omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

const int16_t *refdrv_zetas(void) { return zetas; }
int16_t refdrv_montgomery_reduce(int32_t a) { return montgomery_reduce(a); }
int16_t refdrv_barrett_reduce(int16_t a) { return barrett_reduce(a); }

void refdrv_basemul(int16_t r[2], const int16_t a[2], const int16_t b[2], int16_t zeta)
{
  basemul(r, a, b, zeta);
}

void refdrv_ntt_batch(int16_t *p, size_t npoly, int nthreads)
{
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long i = 0; i < n; i++)
    ntt(p + (size_t)i * 256);
}

void refdrv_invntt_batch(int16_t *p, size_t npoly, int nthreads)
{
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long i = 0; i < n; i++)
    invntt(p + (size_t)i * 256);
}

void refdrv_roundtrip_batch(int16_t *p, size_t npoly, int nthreads)
{
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long i = 0; i < n; i++) {
    ntt(p + (size_t)i * 256);
    invntt(p + (size_t)i * 256);
  }
}

int refdrv_max_threads(void) { return nthr(0); }

/* Whole-poly basemul from the reference's pair basemul with the pairing of
 * ntt()'s output order: pair (4i,4i+1) with zetas[64+i], (4i+2,4i+3) with
 * -zetas[64+i] (pqcrystals poly_basemul_montgomery, upstream of ref/). */
static void poly_basemul_ref(int16_t r[256], const int16_t a[256], const int16_t b[256])
{
  for (int i = 0; i < 64; i++) {
    basemul(&r[4 * i], &a[4 * i], &b[4 * i], zetas[64 + i]);
    basemul(&r[4 * i + 2], &a[4 * i + 2], &b[4 * i + 2], (int16_t)-zetas[64 + i]);
  }
}

void refdrv_basemul_batch(int16_t *r, const int16_t *a, const int16_t *b, size_t npoly,
                          int nthreads)
{
  const long long n = (long long)npoly;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long i = 0; i < n; i++)
    poly_basemul_ref(r + (size_t)i * 256, a + (size_t)i * 256, b + (size_t)i * 256);
}

/* t_i = invntt(reduce(sum_j basemul(A_hat[i][j], ntt(s_j)))) per instance, the
 * accumulation of pqcrystals polyvec_basemul_acc_montgomery (upstream of ref/)
 * built from the reference's own ntt, basemul, barrett_reduce and invntt. */
void refdrv_matvec_batch(int16_t *t, const int16_t *A_hat, const int16_t *s, int k, size_t ninst,
                         int nthreads)
{
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t shat[4][256], tmp[256];
    const int16_t *A = A_hat + (size_t)inst * k * k * 256;
    for (int j = 0; j < k; j++) {
      memcpy(shat[j], s + ((size_t)inst * k + j) * 256, sizeof(shat[j]));
      ntt(shat[j]);
    }
    for (int i = 0; i < k; i++) {
      int16_t *acc = t + ((size_t)inst * k + i) * 256;
      poly_basemul_ref(acc, A + ((size_t)i * k + 0) * 256, shat[0]);
      for (int j = 1; j < k; j++) {
        poly_basemul_ref(tmp, A + ((size_t)i * k + j) * 256, shat[j]);
        for (int c = 0; c < 256; c++)
          acc[c] = (int16_t)(acc[c] + tmp[c]);
      }
      for (int c = 0; c < 256; c++)
        acc[c] = barrett_reduce(acc[c]);
      invntt(acc);
    }
  }
}

/* K-PKE arithmetic cores composed from the reference's ntt, invntt, basemul,
 * montgomery_reduce and barrett_reduce, in the order of the upstream
 * pqcrystals-kyber ref/indcpa.c (poly_ntt = ntt + poly_reduce; basemul_acc =
 * column-0 basemul, int16 adds, poly_reduce; poly_tomont = montgomery_reduce
 * of a * (2^32 mod q)). */
static void rd_reduce(int16_t r[256])
{
  for (int i = 0; i < 256; i++)
    r[i] = barrett_reduce(r[i]);
}

static void rd_poly_ntt(int16_t r[256])
{
  ntt(r);
  rd_reduce(r);
}

static void rd_basemul_acc(int16_t r[256], const int16_t *a, const int16_t *b, int k)
{
  int16_t t[256];
  poly_basemul_ref(r, a, b);
  for (int j = 1; j < k; j++) {
    poly_basemul_ref(t, a + (size_t)j * 256, b + (size_t)j * 256);
    for (int c = 0; c < 256; c++)
      r[c] = (int16_t)(r[c] + t[c]);
  }
  rd_reduce(r);
}

void refdrv_kpke_keygen_batch(int16_t *t_hat, int16_t *s_hat, const int16_t *A_hat,
                              const int16_t *s, const int16_t *e, int k, size_t ninst,
                              int nthreads)
{
  const int16_t f = (int16_t)((1ULL << 32) % 3329);
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t eh[4][256];
    int16_t *sh = s_hat + (size_t)inst * k * 256;
    memcpy(sh, s + (size_t)inst * k * 256, (size_t)k * 256 * sizeof(int16_t));
    for (int j = 0; j < k; j++) {
      rd_poly_ntt(sh + (size_t)j * 256);
      memcpy(eh[j], e + ((size_t)inst * k + j) * 256, sizeof(eh[j]));
      rd_poly_ntt(eh[j]);
    }
    for (int i = 0; i < k; i++) {
      int16_t *pk = t_hat + ((size_t)inst * k + i) * 256;
      rd_basemul_acc(pk, A_hat + ((size_t)inst * k + i) * k * 256, sh, k);
      for (int c = 0; c < 256; c++)
        pk[c] = montgomery_reduce((int32_t)pk[c] * f);
      for (int c = 0; c < 256; c++)
        pk[c] = (int16_t)(pk[c] + eh[i][c]);
      rd_reduce(pk);
    }
  }
}

void refdrv_kpke_encrypt_batch(int16_t *u, int16_t *v, const int16_t *At_hat,
                               const int16_t *t_hat, const int16_t *r, const int16_t *e1,
                               const int16_t *e2, const int16_t *m, int k, size_t ninst,
                               int nthreads)
{
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t rh[4 * 256];
    memcpy(rh, r + (size_t)inst * k * 256, (size_t)k * 256 * sizeof(int16_t));
    for (int j = 0; j < k; j++)
      rd_poly_ntt(rh + (size_t)j * 256);
    for (int i = 0; i < k; i++) {
      int16_t *b = u + ((size_t)inst * k + i) * 256;
      const int16_t *ei = e1 + ((size_t)inst * k + i) * 256;
      rd_basemul_acc(b, At_hat + ((size_t)inst * k + i) * k * 256, rh, k);
      invntt(b);
      for (int c = 0; c < 256; c++)
        b[c] = (int16_t)(b[c] + ei[c]);
      rd_reduce(b);
    }
    int16_t *vv = v + (size_t)inst * 256;
    rd_basemul_acc(vv, t_hat + (size_t)inst * k * 256, rh, k);
    invntt(vv);
    for (int c = 0; c < 256; c++)
      vv[c] = (int16_t)(vv[c] + e2[(size_t)inst * 256 + c]);
    if (m)
      for (int c = 0; c < 256; c++)
        vv[c] = (int16_t)(vv[c] + m[(size_t)inst * 256 + c]);
    rd_reduce(vv);
  }
}

void refdrv_kpke_decrypt_batch(int16_t *w, const int16_t *s_hat, const int16_t *u,
                               const int16_t *v, int k, size_t ninst, int nthreads)
{
  const long long n = (long long)ninst;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
  for (long long inst = 0; inst < n; inst++) {
    int16_t bh[4 * 256];
    memcpy(bh, u + (size_t)inst * k * 256, (size_t)k * 256 * sizeof(int16_t));
    for (int j = 0; j < k; j++)
      rd_poly_ntt(bh + (size_t)j * 256);
    int16_t *mp = w + (size_t)inst * 256;
    rd_basemul_acc(mp, s_hat + (size_t)inst * k * 256, bh, k);
    invntt(mp);
    for (int c = 0; c < 256; c++)
      mp[c] = (int16_t)(v[(size_t)inst * 256 + c] - mp[c]);
    rd_reduce(mp);
  }
}
