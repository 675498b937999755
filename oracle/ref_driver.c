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
