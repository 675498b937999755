/*
 * tkyber_workload.h -- deterministic synthetic inputs (host only).
 *
 * The reference draws its benchmark inputs with glibc rand() seeded with a
 * constant (ntt_bench.cu:941-945), which is not portable.  These generators
 * replace it with a counter-based SplitMix64 stream, so the test suite, the
 * golden-digest script and the benchmark produce identical inputs on any host
 * and any thread count: value i of stream s under seed e is
 * mix64(key(e, s) + (i + 1) * 0x9E3779B97F4A7C15), key(e, s) = mix64(e) ^ s,
 * and a draw in [0, n) is ((v >> 32) * n) >> 32.
 */
#ifndef TKYBER_WORKLOAD_H
#define TKYBER_WORKLOAD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Stream identifiers (any 64-bit value works; these name the ones used). */
#define TK_WL_COEF 0x434f4546ull /* coefficient values      */
#define TK_WL_CAT 0x434154ull    /* per-poly category       */
#define TK_WL_POS 0x504f53ull    /* impulse position        */
#define TK_WL_REQ 0x524551ull    /* ragged request shapes   */
#define TK_WL_GAP 0x474150ull    /* ragged placement gaps   */
#define TK_WL_MAT 0x4d4154ull    /* matrix A_hat            */
#define TK_WL_SEC 0x534543ull    /* secret vector s         */

uint64_t tk_wl_draw(uint64_t seed, uint64_t stream, uint64_t index);
uint32_t tk_wl_below(uint64_t seed, uint64_t stream, uint64_t index, uint32_t n);

/* out[i] = lo + below(n = hi - lo) over stream, for i < n. */
void tk_wl_uniform(int16_t *out, size_t n, uint64_t seed, uint64_t stream, int lo, int hi);

/* Edge-value mix, one category per poly p (below(p, 100)): [0,70) uniform in
 * [-3328, 3328]; [70,80) all zero; [80,90) all 3328; [90,95) all -3328;
 * [95,100) a unit impulse (value 1) at index below(p, 256), zero elsewhere. */
void tk_wl_edge_mix(int16_t *out, size_t npoly, uint64_t seed);

/* Centred binomial eta = 2: (b0 + b1) - (b2 + b3) from four bits of one draw. */
void tk_wl_cbd2(int16_t *out, size_t n, uint64_t seed, uint64_t stream);

typedef struct {
  uint64_t offset; /* first poly of the request inside the pool */
  uint32_t npoly;
  uint16_t k;      /* module rank 2..4 */
  uint16_t kind;   /* 0 = rank-k vector(s), 1 = k x k matrix(es) */
} tk_wl_request;

/* Ragged layout of npoly_total polys: each request is kind 1 (k x k matrices)
 * with probability 1/4, else kind 0 (rank-k vectors), k uniform in {2,3,4};
 * it holds a uniform count of whole objects with at most 4096 polys; the last
 * request is cut so the total is exactly npoly_total.  Requests are placed in
 * order with a gap of below(64) polys before each.  Writes at most max_req
 * entries (reqs may be NULL to count) and returns the number of requests;
 * *pool_polys receives the pool size in polys, gaps included. */
size_t tk_wl_ragged_layout(tk_wl_request *reqs, size_t max_req, size_t npoly_total, uint64_t seed,
                           size_t *pool_polys);

/* FNV-1a 64-bit of the canonical (mod q, in [0,q)) coefficients: a digest of
 * the values, independent of the representative each path emits. */
uint64_t tk_wl_canon_digest(const int16_t *polys, size_t ncoef);

#ifdef __cplusplus
}
#endif
#endif
