/*
 * tkyber.h -- C ABI of the MI355X batched ML-KEM NTT engine.
 *
 * Ring R_q = Z_3329[x]/(x^256 + 1).  A polynomial is int16_t[256] in the
 * reference's coefficient order; a batch is poly-major [npoly][256] in device
 * memory (HBM), 16-byte aligned.  Every entry point takes an explicit stream
 * (a hipStream_t passed as void*, NULL = the legacy default stream), never
 * allocates on the hot path, never exits the process, and returns 0 or a
 * negative tk_status.  All arguments are validated before anything is
 * launched.  Calls from any host thread are allowed; calls on one stream
 * execute in order.
 *
 * Drop-in mapping to the reference (ntt_bench/ in benlwk/TensorKyber):
 *
 *   ntt_batch      replaces ntt(int16_t r[256])              ref/ntt.h:10-11, ref/ntt.c:80-95
 *                  and the batch kernels k_ntt_2tc_tc_cc /   ntt_bench.cu:61-62,181-182
 *                  k_ntt_tc_cc_cp (int16_t* d, int npoly)
 *   invntt_batch   replaces invntt(int16_t r[256])           ref/ntt.h:13-14, ref/ntt.c:106-126
 *                  and k_invntt_2tc_tc_cc / k_invntt_tc_cc_cp ntt_bench.cu:97-98,230-231
 *   basemul_batch  replaces basemul(r, a, b, zeta) applied   ref/ntt.h:16-17, ref/ntt.c:139-146
 *                  to the 128 degree-1 pairs of each poly
 *
 * Output contract (the reference checks mod-q equality after canonicalising
 * both sides, ntt_bench.cu:397-421):
 *   - ntt_batch writes canonical coefficients in [0, q) (the form of the
 *     reference's WMMA forward, ntt_kernels.cuh:448-455);
 *   - invntt_batch writes centred coefficients in [-(q-1)/2, (q-1)/2] (the
 *     form of the reference's WMMA inverse, ntt_kernels.cuh:614-621);
 *   - the _ex variants select the form, including TK_FORM_REF, which yields
 *     the int16 values of ntt()/invntt() themselves (butterfly method only).
 * Inputs may be any int16 values.  The MFMA method computes the exact
 * transform mod q for every input; ref ntt() itself is exact mod q whenever
 * its int16 intermediates do not wrap, which holds for |x| <= 16927 (ntt) and
 * |x| <= 16383 (invntt).
 */
#ifndef TKYBER_H
#define TKYBER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TK_N 256
#define TK_Q 3329

typedef enum {
  TK_OK = 0,
  TK_EINVAL = -22,   /* null / misaligned pointer, zero size, bad enum, bad table */
  TK_ENOMEM = -12,   /* device allocation failed (init / plan creation only)      */
  TK_EDEVICE = -5,   /* a HIP call failed; tk_last_hip_error() has its code        */
  TK_ENOTSUP = -95,  /* combination not provided (e.g. TK_FORM_REF with MFMA)     */
  TK_ENODEV = -19    /* no gfx950 device visible                                  */
} tk_status;

typedef enum {
  TK_METHOD_AUTO = 0,      /* MFMA for canonical/centred forms, butterfly for REF */
  TK_METHOD_BUTTERFLY = 1, /* scalar Cooley-Tukey / Gentleman-Sande kernel        */
  TK_METHOD_MFMA = 2,      /* INT8-MFMA matrix formulation, 7-layer exact         */
  TK_METHOD_MFMA16 = 3,    /* MFMA, v_mfma_i32_16x16x32_i8 tiles (16 polys/wave)  */
  TK_METHOD_MFMA32 = 4     /* MFMA, v_mfma_i32_32x32x32_i8 tiles (32 polys/wave)  */
} tk_method;

typedef enum {
  TK_FORM_DEFAULT = 0,   /* canonical for ntt_batch, centred for invntt_batch     */
  TK_FORM_CANONICAL = 1, /* [0, q)                                                 */
  TK_FORM_CENTERED = 2,  /* [-(q-1)/2, (q-1)/2]                                    */
  TK_FORM_REF = 3        /* int16 values identical to the reference C functions   */
} tk_form;

/* ---- library state --------------------------------------------------- */

/* Builds the constant tables on the host and uploads them to the current
 * device.  Idempotent and thread-safe; every entry point calls it. */
int tk_init(void);
const char *tk_strerror(int status);
/* hipError_t of the last failing HIP call on this thread (0 if none). */
int tk_last_hip_error(void);
const char *tk_version(void);

/* ---- drop-in batch transforms (device pointers) ------------------------ */

int ntt_batch(int16_t *polys, size_t npoly, void *stream);
int invntt_batch(int16_t *polys, size_t npoly, void *stream);
/* r = a o b in the NTT domain; pair (4i,4i+1) uses zetas[64+i] and pair
 * (4i+2,4i+3) uses -zetas[64+i], the pairing of ntt()'s output order.  The
 * default form is TK_FORM_REF: int16 values identical to basemul(). */
int basemul_batch(int16_t *r, const int16_t *a, const int16_t *b, size_t npoly, void *stream);

/* out may equal in (in place) or must not overlap it. */
int ntt_batch_ex(int16_t *out, const int16_t *in, size_t npoly, int method, int form,
                 void *stream);
int invntt_batch_ex(int16_t *out, const int16_t *in, size_t npoly, int method, int form,
                    void *stream);
int basemul_batch_ex(int16_t *r, const int16_t *a, const int16_t *b, size_t npoly, int form,
                     void *stream);

/* ---- ragged batches: one persistent kernel over a request table -------- */

/* One request: npoly polys at base (device pointer), a rank-k module vector
 * (kind 0, npoly a multiple of k) or k x k matrix (kind 1) of ML-KEM-512/768/
 * 1024 (k = 2, 3, 4).  k and kind are carried for callers; every poly of every
 * request is transformed independently. */
typedef struct {
  int16_t *base;
  uint32_t npoly;
  uint16_t k;
  uint16_t kind;
} tk_request;

/* Host-built table of nreq requests; the plan cuts every request into tiles
 * of the kernel's tile height (16 or 32 polys) and copies the table of tiles
 * ({first poly, number of polys}, 16 bytes per tile) to device memory once,
 * picks the kernel shape with the autotuner (or takes method MFMA16 / MFMA32
 * as given) and builds a one-node hipGraph of the persistent kernel's launch
 * (a kernel node, not a stream capture).
 * direction: 0 = forward (canonical out), 1 = inverse (centred out).
 * Transforms are in place.  Plan creation allocates and synchronises; the
 * execute and launch calls do neither.  Requests must not overlap one
 * another (TK_EINVAL). */
typedef struct tk_ragged_plan tk_ragged_plan;
int tk_ragged_plan_create(tk_ragged_plan **plan, const tk_request *table, size_t nreq,
                          int direction, int method, void *stream);
/* Enqueues the plan on stream (NULL = the legacy default stream): one
 * dispatch of the persistent kernel, whatever the number of requests.  (A
 * replay of the plan's one-node graph costs 4-10 us more per launch on
 * ROCm 7; the graph is for composition, below.) */
int tk_ragged_plan_execute(tk_ragged_plan *plan, void *stream);
/* The same as tk_ragged_plan_execute. */
int tk_ragged_plan_launch(tk_ragged_plan *plan, void *stream);
/* The plan's launch as a hipGraph_t (owned by the plan, valid until
 * it is destroyed), for callers that build a graph of their own work and add
 * plans to it as child graph nodes (hipGraphAddChildGraphNode copies it). */
int tk_ragged_plan_graph(tk_ragged_plan *plan, void **graph);
/* Shape the plan runs: method (MFMA16 / MFMA32), polys per wave tile, grid
 * oversubscription (grid = factor x workgroups resident at once), and the
 * total number of polys. */
int tk_ragged_plan_info(const tk_ragged_plan *plan, int *method, int *polys_per_wave,
                        int *oversub, size_t *npoly_total);
int tk_ragged_plan_destroy(tk_ragged_plan *plan);

/* ---- ragged transform without allocation ------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under hipGraph capture, tight serving loops, harnesses that forbid device
 * allocation): the transform of tk_ragged_plan_create + tk_ragged_plan_execute
 * in one call, with no plan.  The request table (host memory, as for a
 * plan) is copied on stream into a fixed device buffer of this library,
 * expanded there into the table of tiles of 16 polys, and the persistent
 * kernel runs over it (MFMA16 tiles, the linear batches' launch shape);
 * forward output canonical, inverse centred, in place.
 *   - Allocates nothing.  tk_init() must have run on the current device
 *     (it sets the buffer up); TK_EINVAL otherwise.
 *   - At most TK_NOALLOC_MAX_REQUESTS requests and TK_NOALLOC_MAX_TILES
 *     tiles (a request takes ceil(npoly / 16)); larger tables give TK_EINVAL.
 *   - The buffer (128 KiB + 2 MiB) is one per device and holds the last
 *     call's contents until the next: calls that overlap on one device
 *     (other streams, other threads) race.
 *   - The table is the source of an asynchronous copy on stream: it must stay
 *     unchanged until the stream has passed this call.
 *   - Requests must not overlap one another; this call does not check it. */
#define TK_NOALLOC_MAX_REQUESTS 8192
#define TK_NOALLOC_MAX_TILES 131072
int tk_ragged_transform_noalloc(const tk_request *table, size_t nreq, int direction,
                                void *stream);

/* ---- fused module matrix-vector product -------------------------------- */

/* For each of ninst instances: t_i = invntt( sum_j A_hat[i][j] o ntt(s_j) ),
 * i, j < k (k in 2..4).  A_hat: [ninst][k][k][256] in the NTT domain, s:
 * [ninst][k][256], t: [ninst][k][256], all device pointers.  One kernel; the
 * NTT of s and the products stay in registers/LDS.  Output is centred. */
int matvec_ntt_fused(int16_t *t, const int16_t *A_hat, const int16_t *s, int k, size_t ninst,
                     void *stream);

/* ---- host-resident batches (PCIe streaming) ---------------------------- */

/* The reference calls ntt()/invntt() per poly on host memory (ref/ntt.h:14-15,
 * ntt_bench.cu:947-956).  A host pipeline runs a host batch on the device:
 * - if in and out each lie wholly inside one pinned allocation
 *   (hipHostMalloc, torch pin_memory), the transform kernel reads and writes
 *   them directly over PCIe (both link directions at once, nothing staged);
 * - otherwise the batch streams through the device in chunks of chunk_polys:
 *   each chunk is copied in, transformed in place and copied out on its own
 *   stream, up to depth chunks in flight; pageable buffers go through the
 *   HIP runtime's staging (slower).
 * Creation allocates depth pinned staging buffers of chunk_polys (larger
 * chunks stage faster: 65536 polys, 32 MiB, reach about 30 GB/s each way)
 * and streams on the current device; run blocks until out is written, on the
 * creating device only. */
#define TK_HOST_MAX_DEPTH 4
#define TK_HOST_DEFAULT_CHUNK 65536 /* polys (32 MiB) per chunk */
#define TK_HOST_DEFAULT_DEPTH 2
typedef struct tk_host_pipeline tk_host_pipeline;
int tk_host_pipeline_create(tk_host_pipeline **pipeline, size_t chunk_polys, int depth);
/* direction 0: forward (default form canonical), 1: inverse (centred);
 * method and form as for ntt_batch_ex; out may equal in (no partial
 * overlap).  in and out are host memory: device memory gives TK_EINVAL.
 * One run at a time per pipeline (others wait). */
int tk_host_pipeline_run(tk_host_pipeline *pipeline, int16_t *out, const int16_t *in,
                         size_t npoly, int direction, int method, int form);
int tk_host_pipeline_destroy(tk_host_pipeline *pipeline);

/* The reference's loop `for (i..n) ntt(poly[i])` on a host array, in place,
 * through a per-device default pipeline (TK_HOST_DEFAULT_CHUNK,
 * TK_HOST_DEFAULT_DEPTH), with the default forms (canonical forward,
 * centred inverse). */
int ntt_batch_host(int16_t *polys, size_t npoly);
int invntt_batch_host(int16_t *polys, size_t npoly);

/* ---- K-PKE arithmetic cores -------------------------------------------- */

/* The arithmetic of ML-KEM's K-PKE around the NTT path, one fused kernel
 * each, composed as the pqcrystals-kyber reference composes it (ref/indcpa.c,
 * ref/polyvec.c, ref/poly.c; upstream of ntt_bench/ref/, which vendors only
 * its ntt.c and reduce.c).  Every output ends in that reference's
 * poly_reduce (barrett_reduce, ref/reduce.c:35-42) and equals its raw int16
 * values: centred representatives in [-(q-1)/2, (q-1)/2].  Sampling,
 * hashing, matrix expansion and (de)compression stay with the caller.
 * Shapes per instance: matrices [k][k][256] (row i = polys i*k .. i*k+k-1),
 * vectors [k][256], single polys [256]; instances back to back; k in 2..4;
 * device pointers, 16-byte aligned.  An output may alias exactly (same
 * address) only the input named beside it; other overlaps are TK_EINVAL. */

/* Key generation (indcpa_keypair_derand): s_hat = poly_ntt(s) (may alias s);
 * t_hat_i = reduce(tomont(sum_j A_hat[i][j] o s_hat_j) + poly_ntt(e_i)) (may
 * alias e), both in the NTT domain. */
int kpke_keygen_core(int16_t *t_hat, int16_t *s_hat, const int16_t *A_hat, const int16_t *s,
                     const int16_t *e, int k, size_t ninst, void *stream);

/* Encryption (indcpa_enc), with At_hat the transposed matrix as gen_at
 * yields it: r_hat = poly_ntt(r); u_i = reduce(invntt(sum_j At_hat[i][j] o
 * r_hat_j) + e1_i) (may alias e1); v = reduce(invntt(sum_j t_hat_j o r_hat_j)
 * + e2 + m) (may alias e2).  m is the message poly (poly_frommsg) or NULL.
 * The adds are exact mod q: equal to the reference's int16 poly_add wherever
 * that does not overflow, which holds for every ML-KEM input (noise of at
 * most eta = 3, message coefficients 0 or (q+1)/2). */
int kpke_encrypt_core(int16_t *u, int16_t *v, const int16_t *At_hat, const int16_t *t_hat,
                      const int16_t *r, const int16_t *e1, const int16_t *e2, const int16_t *m,
                      int k, size_t ninst, void *stream);

/* Decryption (indcpa_dec): w = reduce(v - invntt(sum_j s_hat_j o
 * poly_ntt(u_j))) (may alias v).  The subtraction is exact mod q: equal
 * to the reference's int16 poly_sub wherever that does not overflow (always
 * for decompressed v in [0, q)). */
int kpke_decrypt_core(int16_t *w, const int16_t *s_hat, const int16_t *u, const int16_t *v, int k,
                      size_t ninst, void *stream);

/* ---- rank-aware scheduling of mixed fused requests --------------------- */

/* A serving batch mixes ML-KEM-512/768/1024 (module ranks k = 2/3/4,
 * ref/params.h:10-18) and operations.  A request is a contiguous batch of
 * ninst instances of one operation at one rank, with the buffers of the
 * matching single call (which the request's roles map onto):
 *   TK_FUSED_MATVEC  (matvec_ntt_fused):  out0 = t, mat = A_hat, x = s
 *   TK_FUSED_KEYGEN  (kpke_keygen_core):  out0 = t_hat, out1 = s_hat,
 *                                         mat = A_hat, x = s, add0 = e
 *   TK_FUSED_ENCRYPT (kpke_encrypt_core): out0 = u, out1 = v, mat = At_hat,
 *                                         mat2 = t_hat, x = r, add0 = e1,
 *                                         add1 = e2, add2 = m (or NULL)
 *   TK_FUSED_DECRYPT (kpke_decrypt_core): out0 = w, mat = s_hat, x = u,
 *                                         add0 = v
 * Unused roles are NULL.  Each request obeys the rules of its single call;
 * different requests' outputs must not overlap any other request's buffers.
 * The plan groups the instances by (operation, rank) -- one kernel per group
 * through a device table of per-instance pointers -- and records the groups,
 * largest first, as parallel branches of one hipGraph.  Creation validates,
 * allocates and synchronises; execute replays the graph (NULL stream = the
 * legacy default stream), launch issues the group kernels in order. */
#define TK_FUSED_MATVEC 0
#define TK_FUSED_KEYGEN 1
#define TK_FUSED_ENCRYPT 2
#define TK_FUSED_DECRYPT 3
typedef struct {
  int mode;
  int k;
  size_t ninst;
  int16_t *out0;
  int16_t *out1;
  const int16_t *mat;
  const int16_t *mat2;
  const int16_t *x;
  const int16_t *add0;
  const int16_t *add1;
  const int16_t *add2;
} tk_fused_request;
typedef struct tk_fused_plan tk_fused_plan;
int tk_fused_plan_create(tk_fused_plan **plan, const tk_fused_request *reqs, size_t nreq);
int tk_fused_plan_execute(tk_fused_plan *plan, void *stream);
int tk_fused_plan_launch(tk_fused_plan *plan, void *stream);
/* Number of (operation, rank) groups and of instances in the plan. */
int tk_fused_plan_info(const tk_fused_plan *plan, int *ngroups, size_t *ninst_total);
int tk_fused_plan_destroy(tk_fused_plan *plan);

/* ---- packed small-coefficient input ------------------------------------ */

/* The packed form of the reference's TC-PACK path (ntt_bench.cu:278-391,
 * 740-773): a poly of small coefficients takes bits-bit two's-complement
 * fields, 8/bits per byte, coefficient i in byte i/(8/bits) at bit
 * bits*(i%(8/bits)), little-endian within the byte (pack_small,
 * ntt_bench.cu:749-757).  bits 2 (ternary), 3 (CBD eta = 2 or 3) or 4; a poly
 * takes tk_packed_poly_bytes(bits) = 64, 128 or 128 bytes, polys back to back.
 * Values outside [-2^(bits-1), 2^(bits-1)) keep their low bits, as in
 * pack_small. */
size_t tk_packed_poly_bytes(int bits); /* 0 for an unsupported width */

/* Packs npoly int16 polys (device) into the packed form (device, 16-byte
 * aligned, not overlapping in). */
int pack_small_batch(uint8_t *packed, const int16_t *in, size_t npoly, int bits, void *stream);

/* Forward NTT of packed polys (k_ntt_tc_pack, ntt_bench.cu:352-391): unpacked
 * while staged, then transformed as ntt_batch_ex does.  method AUTO, MFMA,
 * MFMA16 or MFMA32 (BUTTERFLY: TK_ENOTSUP); form DEFAULT/CANONICAL or
 * CENTERED.  out must not overlap packed. */
int ntt_batch_packed(int16_t *out, const uint8_t *packed, size_t npoly, int bits, int method,
                     int form, void *stream);

/* matvec_ntt_fused with s in the packed form: s_packed is
 * [ninst][k][tk_packed_poly_bytes(bits)].  t must not overlap A_hat or
 * s_packed.  Output is centred. */
int matvec_ntt_fused_packed(int16_t *t, const int16_t *A_hat, const uint8_t *s_packed, int k,
                            size_t ninst, int bits, void *stream);

/* ---- autotuner ---------------------------------------------------------- */

/* Times each MFMA tile shape x waves per workgroup (2, 4) x grid
 * oversubscription (4, 8, 12 times the workgroups resident at once) of the
 * ragged-plan kernel on scratch data of npoly polys (at most 2^20) laid out
 * as one request, on the current device, in interleaved rounds, and returns
 * the fastest -- unless it beats the linear batches' shape (MFMA16, 4 waves,
 * 8x) by less than 5%, which is then kept: method (TK_METHOD_MFMA16 or
 * TK_METHOD_MFMA32), waves per workgroup and the oversubscription factor.
 * Plans with method TK_METHOD_AUTO run the same search on their own
 * requests.  Results are cached per device, direction, size octave and
 * request-size octave for the life of the process.  Allocates and
 * synchronises. */
int tk_autotune(size_t npoly, int direction, int *best_method, int *best_waves_per_wg,
                int *best_oversub, void *stream);

/* ---- introspection (host only, no device needed) ----------------------- */

/* Copies the host-built tables: zetas[128] (Montgomery form, centred, as the
 * reference table), and the plain mod-q stage matrices: fwd1[16*16]
 * (layers 1-4, [row][m]), fwd2[16*16*16] (layers 5-7, [block][c][j]),
 * inv2[16*16*16] (inverse layers 7-5, [block][c][j]) and inv1[16*16]
 * (inverse layers 4-1 with the final 1441 scaling).  Any pointer may be NULL. */
int tk_host_tables(int16_t *zetas, int16_t *fwd1, int16_t *fwd2, int16_t *inv2, int16_t *inv1);

/* Copies the packed MFMA A operands uploaded for shape 16 (16x16x32 tiles,
 * ops = [nmat][2 limbs][64 lanes][2 dwords]) or 32 (32x32x32 tiles,
 * ops = [nmat][64 lanes][4 dwords]) of stage 0 (forward layers 1-4, nmat 1),
 * 1 (forward layers 5-7, nmat 16), 2 (inverse layers 7-5, nmat 16) or 3
 * (inverse layers 4-1, nmat 1), and the row biases bias = [nmat][16].
 * Stages 4, 5, 6 (shape 16 only, nmat 1): forward layers 1-4 for packed
 * 2-, 3-, 4-bit input, whose data digits are the unsigned field f and its
 * sign bit s (x = f - 2^bits s). */
int tk_host_mfma_operands(int shape, int stage, uint32_t *ops, int32_t *bias);

#ifdef __cplusplus
}
#endif
#endif
