/*
 * kyber_oracle.h -- CPU restatement of the reference ML-KEM NTT path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by the CPU-baseline leg of
 * bench.py, never by the product library (tensorkyber_amd), which fails
 * loudly when its HIP build is missing instead of falling back to this code.
 *
 * Every function restates the reference operation for operation, including
 * the int16 narrowing of C assignments and the arithmetic right shift of
 * negative int32 values (gcc semantics), so its outputs are raw-int16
 * identical to the reference, not merely congruent mod q.
 *
 *   reference file                              restated here
 *   ntt_bench/ref/reduce.c:16-23  montgomery_reduce  oracle_montgomery_reduce
 *   ntt_bench/ref/reduce.c:35-42  barrett_reduce     oracle_barrett_reduce
 *   ntt_bench/ref/ntt.c:68-70     fqmul              oracle_fqmul
 *   ntt_bench/ref/ntt.c:39-56     zetas[128]         oracle_zetas (closed form,
 *                                                    generator ntt.c:6-37)
 *   ntt_bench/ref/ntt.c:80-95     ntt                oracle_ntt
 *   ntt_bench/ref/ntt.c:106-126   invntt             oracle_invntt
 *   ntt_bench/ref/ntt.c:139-146   basemul            oracle_basemul
 *   pqcrystals-kyber ref/indcpa.c  (upstream)        oracle_kpke_*_batch
 *   ntt_bench/ntt_bench.cu:749-757 pack_small        oracle_pack_small
 *   ntt_bench/ntt_bench.cu:325-348 pack_unpack       oracle_unpack_small
 *
 * Parity pins (tests/test_oracle.py): raw equality with the compiled
 * reference (oracle/_ref) and with the committed golden fixtures
 * (tests/golden/), exhaustive int16 / full-precondition checks of the
 * reductions against their closed forms, brute-force O(N^2) evaluation of the
 * negacyclic NTT mod q, invntt(ntt(x)) == 2^16 x mod q, and the negacyclic
 * schoolbook product for the whole-poly basemul.
 */
#ifndef KYBER_ORACLE_H
#define KYBER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_N 256
#define ORACLE_Q 3329

int16_t oracle_montgomery_reduce(int32_t a);
int16_t oracle_barrett_reduce(int16_t a);
int16_t oracle_fqmul(int16_t a, int16_t b);

/* zetas[i] = centred(17^brv7(i) * 2^16 mod q); returns a pointer to 128 entries. */
const int16_t *oracle_zetas(void);

void oracle_ntt(int16_t r[256]);
void oracle_invntt(int16_t r[256]);
void oracle_basemul(int16_t r[2], const int16_t a[2], const int16_t b[2], int16_t zeta);

/* Whole-poly product in the NTT domain: pair (4i,4i+1) uses zetas[64+i], pair
 * (4i+2,4i+3) uses -zetas[64+i] (the pairing of ntt()'s bit-reversed output
 * order, as in pqcrystals poly_basemul_montgomery, upstream of ref/). */
void oracle_poly_basemul(int16_t r[256], const int16_t a[256], const int16_t b[256]);

/* Batched drivers over [npoly][256] poly-major arrays.  nthreads <= 0 uses
 * every OpenMP thread. */
void oracle_ntt_batch(int16_t *polys, size_t npoly, int nthreads);
void oracle_invntt_batch(int16_t *polys, size_t npoly, int nthreads);
void oracle_basemul_batch(int16_t *r, const int16_t *a, const int16_t *b, size_t npoly,
                          int nthreads);
/* NTT then inverse NTT of each poly in place (the round trip the benchmark times). */
void oracle_roundtrip_batch(int16_t *polys, size_t npoly, int nthreads);

/* Module matrix-vector product t = invntt(A_hat o ntt(s)) per instance.
 * A_hat: [ninst][k][k][256] (NTT domain), s: [ninst][k][256], t: [ninst][k][256].
 * Accumulation follows pqcrystals polyvec_basemul_acc_montgomery (upstream of
 * ref/): basemul of column 0, int16 poly_add of the other columns, Barrett
 * reduce of every coefficient, then invntt. */
void oracle_matvec_batch(int16_t *t, const int16_t *A_hat, const int16_t *s, int k,
                         size_t ninst, int nthreads);

/* K-PKE arithmetic cores of ML-KEM, composed as the pqcrystals-kyber
 * reference composes them in ref/indcpa.c (upstream of ref/, not vendored):
 * poly_ntt = ntt then poly_reduce; polyvec_basemul_acc_montgomery = basemul
 * of column 0, int16 adds of the others, poly_reduce; poly_tomont = fqmul by
 * 2^32 mod q; poly_reduce = barrett_reduce of every coefficient; poly_add /
 * poly_sub = int16 adds.  Every output ends in poly_reduce (centred).
 *   keygen:  s_hat = poly_ntt(s); t_hat_i = reduce(tomont(A_hat_i . s_hat) +
 *            poly_ntt(e_i))                       (indcpa_keypair_derand)
 *   encrypt: r_hat = poly_ntt(r); u_i = reduce(invntt(At_hat_i . r_hat) +
 *            e1_i); v = reduce(invntt(t_hat . r_hat) + e2 + m)  (indcpa_enc)
 *   decrypt: w = reduce(v - invntt(s_hat . poly_ntt(u)))          (indcpa_dec)
 * Shapes: matrices [ninst][k][k][256] (row i = polys i*k .. i*k+k-1),
 * vectors [ninst][k][256], single polys [ninst][256].  m may be NULL. */
void oracle_kpke_keygen_batch(int16_t *t_hat, int16_t *s_hat, const int16_t *A_hat,
                              const int16_t *s, const int16_t *e, int k, size_t ninst,
                              int nthreads);
void oracle_kpke_encrypt_batch(int16_t *u, int16_t *v, const int16_t *At_hat,
                               const int16_t *t_hat, const int16_t *r, const int16_t *e1,
                               const int16_t *e2, const int16_t *m, int k, size_t ninst,
                               int nthreads);
void oracle_kpke_decrypt_batch(int16_t *w, const int16_t *s_hat, const int16_t *u,
                               const int16_t *v, int k, size_t ninst, int nthreads);

/* Packed small-coefficient polys (the TC-PACK input format): bits-bit
 * two's-complement fields, per = 8/bits per byte, coefficient i in byte
 * i/per at bit bits*(i%per).  A poly takes 256/per bytes (bits 2: 64;
 * bits 3, 4: 128).  pack keeps the low bits of each value; unpack
 * sign-extends each field.  Batched over npoly polys back to back. */
size_t oracle_packed_poly_bytes(int bits);
void oracle_pack_small(uint8_t *packed, const int16_t *polys, size_t npoly, int bits);
void oracle_unpack_small(int16_t *polys, const uint8_t *packed, size_t npoly, int bits);

/* Canonical representative in [0, q). */
static inline int16_t oracle_canon(int16_t v)
{
  int r = v % ORACLE_Q;
  return (int16_t)(r < 0 ? r + ORACLE_Q : r);
}

int oracle_max_threads(void);

/* Whole-domain invariant checks of the reductions; returns the violation count. */
long long oracle_check_reductions(void);

#ifdef __cplusplus
}
#endif
#endif
