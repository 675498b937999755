// Chosen method: the seven NTT layers as two exact INT8 matrix products on
// the CDNA4 matrix cores.
//
// A poly is x[m][c] = r[16m + c].  Forward: layers 1-4 are one 16x16 matrix
// over m (the same for every c), layers 5-7 are one 16x16 matrix per block m
// over c.  Inverse: the per-block matrices first, then one matrix over m that
// also carries the final 1441 scaling (ref/ntt.c:124-125).  Every constant is
// split into signed 8-bit limbs, the data into two signed byte digits (see
// DeviceTables in tk_internal.h), the products accumulate exactly in int32,
// and each output is reduced once with a 24-bit Barrett step.  The result is
// exact mod q for every int16 input.
//
// Data is the MFMA B operand (one poly per output column, so a wave works on
// 16 or 32 polys at once) and the constants are the A operand, resident in
// VGPRs for the life of the wave.  Each wave owns an LDS slab holding its
// tile; the two passes read and write only lane-private LDS words, so the only
// cross-lane hand-offs are tile load -> pass 1 -> pass 2 -> tile store.
#include <type_traits>

#include "tk_launch.h"
#include "tk_mfma_core.h"

namespace tk {
namespace {

// ---------------------------------------------------------------------------
// Kernels.  INV selects the inverse; FORM the output form of the last pass.
// Grid-stride over tiles of P polys per wave; WAVES waves per workgroup.
// ---------------------------------------------------------------------------
template <int FORM, bool INV>
__device__ __forceinline__ int32_t bias_adjust()
{
  // Forward tables carry no centring offset, inverse final tables do.
  if (!INV)
    return FORM == kFormCentred ? kCentre : 0;
  return FORM == kFormCentred ? 0 : -kCentre;
}

// Per-tile inputs and outputs, wave-uniform: where tile `tile`'s polys are
// read and written and how many of its P polys there are.  Linear and packed
// tiles follow from the tile index; ragged tiles come from the plan's tile
// table, one scalar load per tile (the table is read-only for the kernel's
// life, so it is read through the constant address space).
struct TileIo {
  const int16_t *in;
  int16_t *out;
  uint32_t npoly;
};

template <int P>
__device__ __forceinline__ size_t tile_count(const LinearTiles &io)
{
  return (io.npoly + P - 1) / P;
}

template <int P>
__device__ __forceinline__ TileIo tile_io(const LinearTiles &io, size_t tile)
{
  const size_t left = io.npoly - tile * P;
  return TileIo{io.in + tile * P * kN, io.out + tile * P * kN,
                (uint32_t)(left < (size_t)P ? left : (size_t)P)};
}

template <int P>
__device__ __forceinline__ size_t tile_count(const RaggedTiles &io)
{
  return io.ntiles;
}

template <int P>
__device__ __forceinline__ TileIo tile_io(const RaggedTiles &io, size_t tile)
{
  typedef const __attribute__((address_space(4))) TileRef *ConstTiles;
  const ConstTiles t = (ConstTiles)io.tiles + tile;
  int16_t *const base = t->base;
  return TileIo{base, base, t->npoly};
}

template <int P, int BITS>
__device__ __forceinline__ size_t tile_count(const PackedTiles<BITS> &io)
{
  return (io.npoly + P - 1) / P;
}

template <int P, int BITS>
__device__ __forceinline__ TileIo tile_io(const PackedTiles<BITS> &io, size_t tile)
{
  const size_t left = io.npoly - tile * P;
  return TileIo{nullptr, io.out + tile * P * kN, (uint32_t)(left < (size_t)P ? left : (size_t)P)};
}

// Staging registers: packed tiles are small (4-16 dwords per lane), so the
// next tile's loads are issued into registers before the current tile's
// stores.  The wait for those loads then does not include the stores'
// completion (vmcnt counts both, in issue order).  An int16 tile (8 KB per
// wave) held that way costs a wave of occupancy, so it is staged directly.
template <int P, class Tiles>
struct StageRegs {
};

template <int P, int BITS>
struct StageRegs<P, PackedTiles<BITS>> {
  uint32_t v[packed_regs<P, BITS>()];
};

template <int P, class Tiles>
__device__ __forceinline__ void stage_load(const Tiles &, size_t, StageRegs<P, Tiles> &, int)
{
}

// Packed tiles of 16 polys (MFMA16) are loaded in the lane layout of the
// direct first pass (lane (g, t): coefficients 64g .. 64g+63 of poly t) and
// never unpacked; 32-poly tiles (MFMA32) are unpacked into the slab.
template <int P, int BITS>
__device__ __forceinline__ void stage_load(const PackedTiles<BITS> &io, size_t tile,
                                           StageRegs<P, PackedTiles<BITS>> &r, int lane)
{
  constexpr size_t PPB = packed_poly_bytes<BITS>();
  auto src = [&](int pp) -> const uint8_t * {
    const size_t p = tile * P + pp;
    return p < io.npoly ? io.in + p * PPB : nullptr;
  };
  if constexpr (P == 16) {
    static_assert(packed_regs<P, BITS>() == packed_lane_dwords<BITS>(), "one layout per lane");
    packed_load_lanes<BITS>(r.v, src, lane);
  } else {
    packed_load<P, BITS>(r.v, src, lane);
  }
}

// Stage the tile into the slab L: int16 tiles (linear and ragged) through a
// buffer descriptor over the tile's polys, packed polys from the staging
// registers.
template <int P, class Tiles>
__device__ __forceinline__ void stage_in(const Tiles &, const TileIo &t, int16_t *L,
                                         const StageRegs<P, Tiles> &, int lane)
{
  tile_in_rsrc<P>(L, span_rsrc(t.in, t.npoly * 2 * kN), lane);
}

template <int P, int BITS>
__device__ __forceinline__ void stage_in(const PackedTiles<BITS> &, const TileIo &, int16_t *L,
                                         const StageRegs<P, PackedTiles<BITS>> &r, int lane)
{
  if constexpr (P != 16)
    packed_park<P, BITS>(L, r.v, lane);
}

// Write the slab to the tile's output polys, with cache policy POL.
template <int P, int POL>
__device__ __forceinline__ void stage_out(const TileIo &t, const int16_t *L, int lane)
{
  tile_out_rsrc<P, POL>(L, span_rsrc(t.out, t.npoly * 2 * kN), lane);
}

// Bits per coefficient of packed tiles, 0 for int16 tiles.
template <class Tiles>
struct PackedBits : std::integral_constant<int, 0> {
};
template <int BITS>
struct PackedBits<PackedTiles<BITS>> : std::integral_constant<int, BITS> {
};

// Four waves per SIMD (at most 128 VGPRs), which LDS allows too (an
// 8,448-byte slab per wave).  SWAP: int16 tiles swap (below); without it each tile is
// staged in after the previous one is stored, which the launcher picks when
// every wave has one tile: the swap's prologue then has nothing to overlap
// (262,144 polys: 47-48 against 50-51 us forward).
//
// QUEUED: the waves of a workgroup take their tiles from a counter in LDS
// instead of striding over them: workgroup b's k-th tile is k * grid + b,
// handed to whichever of its waves asks next.  With one workgroup per CU
// (the launcher's shape for packed input) the waves of a SIMD then share
// its tiles.  Striding, each wave had the same number of tiles
// while the SIMD issues for the oldest ready wave first: the four finished
// in a staircase (2^20 2-bit polys: at about 60, 80, 100 and 118 us), and
// the last third of the kernel ran on fewer waves than the SIMD holds.
//
// KEEP: the output is stored with the default cache policy instead of
// nontemporally, so that it stays in the Infinity Cache (keep_output).
template <int FORM, bool INV, int WAVES, class Tiles, bool SWAP, bool QUEUED = false,
          bool KEEP = false>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(4))) void k_mfma16(Tiles io,
                                                        const DeviceTables *__restrict__ tab)
{
  constexpr int P = 16;
  __shared__ __align__(16) int16_t s_poly[WAVES][P * kStride];
  __shared__ __align__(16) int32_t s_bias[16][16];
  __shared__ uint32_t s_next;  // QUEUED: the workgroup's next tile number
  if (QUEUED && threadIdx.x == 0)
    s_next = 0;
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), g = lane >> 4, t = lane & 15;

  const int32_t(*bias_c)[16] = INV ? tab->bias_inv2 : tab->bias_fwd2;
  for (int i = threadIdx.x; i < 256; i += 64 * WAVES)
    s_bias[i >> 4][i & 15] = bias_c[i >> 4][i & 15] + (INV ? 0 : bias_adjust<FORM, INV>());
  __syncthreads();

  // Packed input runs its first pass on (field, sign bit) digits.
  constexpr int PB = PackedBits<Tiles>::value;
  const auto &m1 = INV ? tab->m16_inv1[0] : (PB ? tab->m16_fwd1pk[PB ? PB - 2 : 0] : tab->m16_fwd1[0]);
  const auto &m2 = INV ? tab->m16_inv2 : tab->m16_fwd2;
  const long a1lo = *reinterpret_cast<const long *>(m1[0][lane]);
  const long a1hi = *reinterpret_cast<const long *>(m1[1][lane]);
  const int32_t *b1p = INV ? tab->bias_inv1[0] : (PB ? tab->bias_fwd1pk[0] : tab->bias_fwd1[0]);
  const int32_t b1adj = INV ? bias_adjust<FORM, INV>() : 0;
  const v4i bias1 = {b1p[4 * g] + b1adj, b1p[4 * g + 1] + b1adj, b1p[4 * g + 2] + b1adj,
                     b1p[4 * g + 3] + b1adj};
  long a2lo[16], a2hi[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    a2lo[m] = *reinterpret_cast<const long *>(m2[m][0][lane]);
    a2hi[m] = *reinterpret_cast<const long *>(m2[m][1][lane]);
  }

  int16_t *L = s_poly[w];
  int16_t *Pp = L + t * kStride;
  // Grid-stride over tiles.  Latency is hidden by the other resident waves
  // (three or four per SIMD) and by launching several times more workgroups
  // than are resident, which also balances load across CUs.  The next
  // tile's description (for ragged tiles a scalar load of the table) is
  // issued at the top of each iteration, a tile ahead of its use.
  //
  // Vector memory operations complete in issue order, so no wait for a load
  // may fall shortly after the previous tile's stores: it would wait for
  // them as well (on MI355X a write is acknowledged microseconds after its
  // store issues).  The next tile's loads therefore go out before the
  // current tile's stores: int16 tiles swap (tile_swap_rsrc16: loads,
  // stores, then the waits for the loads), packed tiles load into the
  // staging registers as soon as the first pass has consumed them.  Both
  // measured faster than loading after the stores: 180-182 against 183 us
  // per 2^20-poly int16 pass, 126-130 against 138-140 us for the 2-bit
  // packed forward (before its tile queue).  vm_drain keeps the loop's waits from covering the
  // prologue's loads.  Without SWAP (one tile per wave) the int16 tile is
  // staged in at the top of the loop and the prologue is not drained, so the
  // matrices' loads stay in flight through the tile's first pass.
  const size_t ntiles = tile_count<P>(io);
  const size_t stride = (size_t)gridDim.x * WAVES;
  auto take = [&]() -> size_t {
    uint32_t k = 0;
    if (lane == 0)
      k = __hip_atomic_fetch_add(&s_next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return (size_t)__builtin_amdgcn_readfirstlane(k) * gridDim.x + blockIdx.x;
  };
  size_t tile = QUEUED ? take() : (size_t)blockIdx.x * WAVES + w;
  TileIo cur = tile_io<P>(io, tile < ntiles ? tile : 0);
  StageRegs<P, Tiles> sr;
  constexpr bool kSwap = PB == 0 && SWAP;
  constexpr int kOut = KEEP ? kBufferKeep : kBufferNT;
  stage_load<P>(io, tile, sr, lane);
  if constexpr (kSwap)
    if (tile < ntiles)
      stage_in<P>(io, cur, L, sr, lane);
  if constexpr (PB != 0 || SWAP)
    vm_drain();
  for (size_t next; tile < ntiles; tile = next) {
    next = QUEUED ? take() : tile + stride;
    const bool more = next < ntiles;
    const TileIo nxt = tile_io<P>(io, more ? next : tile);
    if constexpr (!kSwap)
      stage_in<P>(io, cur, L, sr, lane);
    wave_sync();
    if (!INV) {
      if constexpr (PB != 0) {
        pass_m16_packed<PB>(Pp, g, sr.v, a1lo, a1hi, bias1);
        stage_load<P>(io, next, sr, lane);
      } else {
        pass_m16<kFormInter>(Pp, g, a1lo, a1hi, bias1);
      }
      wave_sync();
      pass_c16<FORM>(Pp, g, [&](int m) { return a2lo[m]; }, [&](int m) { return a2hi[m]; },
                     s_bias);
    } else {
      pass_c16<kFormInter>(Pp, g, [&](int m) { return a2lo[m]; },
                           [&](int m) { return a2hi[m]; }, s_bias);
      wave_sync();
      pass_m16<FORM>(Pp, g, a1lo, a1hi, bias1);
    }
    wave_sync();
    if (kSwap && more)
      tile_swap_rsrc16<kOut>(L, span_rsrc(cur.out, cur.npoly * 2 * kN),
                             span_rsrc(nxt.in, nxt.npoly * 2 * kN), lane);
    else
      stage_out<P, kOut>(cur, L, lane);
    wave_sync();
    cur = nxt;
  }
}

// Two waves per SIMD (at most 256 VGPRs), as for k_mfma16.
template <int FORM, bool INV, int WAVES, class Tiles, bool KEEP = false>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(2))) void k_mfma32(Tiles io,
                                                        const DeviceTables *__restrict__ tab)
{
  constexpr int P = 32;
  __shared__ __align__(16) int16_t s_poly[WAVES][P * kStride];
  __shared__ __align__(16) int32_t s_bias[16][16];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), h = lane >> 5, t = lane & 31;

  const int32_t(*bias_c)[16] = INV ? tab->bias_inv2 : tab->bias_fwd2;
  for (int i = threadIdx.x; i < 256; i += 64 * WAVES)
    s_bias[i >> 4][i & 15] = bias_c[i >> 4][i & 15] + (INV ? 0 : bias_adjust<FORM, INV>());
  __syncthreads();

  const auto &m1 = INV ? tab->m32_inv1[0] : tab->m32_fwd1[0];
  const auto &m2 = INV ? tab->m32_inv2 : tab->m32_fwd2;
  const v4i a1 = *reinterpret_cast<const v4i *>(m1[lane]);
  const int32_t *b1p = INV ? tab->bias_inv1[0] : tab->bias_fwd1[0];
  const int32_t b1adj = INV ? bias_adjust<FORM, INV>() : 0;
  const v4i bias1a = {b1p[4 * h] + b1adj, b1p[4 * h + 1] + b1adj, b1p[4 * h + 2] + b1adj,
                      b1p[4 * h + 3] + b1adj};
  const v4i bias1b = {b1p[8 + 4 * h] + b1adj, b1p[8 + 4 * h + 1] + b1adj,
                      b1p[8 + 4 * h + 2] + b1adj, b1p[8 + 4 * h + 3] + b1adj};
  v4i a2[16];
#pragma unroll
  for (int m = 0; m < 16; m++)
    a2[m] = *reinterpret_cast<const v4i *>(m2[m][lane]);

  int16_t *L = s_poly[w];
  int16_t *Pp = L + t * kStride;
  // 32 polys per wave: no register budget left for a prefetch buffer.
  const size_t ntiles = tile_count<P>(io);
  const size_t stride = (size_t)gridDim.x * WAVES;
  size_t tile = (size_t)blockIdx.x * WAVES + w;
  TileIo cur = tile_io<P>(io, tile < ntiles ? tile : 0);
  StageRegs<P, Tiles> sr;
  stage_load<P>(io, tile, sr, lane);
  for (; tile < ntiles; tile += stride) {
    const TileIo nxt = tile_io<P>(io, tile + stride < ntiles ? tile + stride : tile);
    stage_in<P>(io, cur, L, sr, lane);
    wave_sync();
    if (!INV) {
      pass_m32<kFormInter>(Pp, h, a1, bias1a, bias1b);
      wave_sync();
      pass_c32<FORM>(Pp, h, [&](int m) { return a2[m]; }, s_bias);
    } else {
      pass_c32<kFormInter>(Pp, h, [&](int m) { return a2[m]; }, s_bias);
      wave_sync();
      pass_m32<FORM>(Pp, h, a1, bias1a, bias1b);
    }
    wave_sync();
    stage_load<P>(io, tile + stride, sr, lane);
    stage_out<P, KEEP ? kBufferKeep : kBufferNT>(cur, L, lane);
    wave_sync();
    cur = nxt;
  }
}

// Waves of a queued workgroup (one per CU): four per SIMD.  With the packed
// loads kept in L2 (packed_load_lanes), 2^20 polys take 104-105 us at 2 bits
// and 108 us at 3 bits, against 107-108 and 110 us with 12 waves and 120-124
// with 8.  (With nontemporal loads 12 waves had been best: 116-118 and 130
// us, against 118-119 and 132-133 with 16.)
constexpr int kQueuedWaves = 16;

inline size_t host_tile_count(const LinearTiles &io, int P) { return (io.npoly + P - 1) / P; }
inline size_t host_tile_count(const RaggedTiles &io, int) { return io.ntiles; }
template <int BITS>
inline size_t host_tile_count(const PackedTiles<BITS> &io, int P)
{
  return (io.npoly + P - 1) / P;
}

template <class K, class Tiles>
int launch_persistent(K kernel, int waves, int P, int ncu, int oversub, const Tiles &io,
                      const DeviceTables *tab, const LaunchSink &s)
{
  const int block = 64 * waves;
  const int per_cu = resident_per_cu((const void *)kernel, block) * (oversub > 0 ? oversub : 1);
  const size_t ntiles = host_tile_count(io, P);
  const size_t need = (ntiles + waves - 1) / waves;
  const size_t cap = (size_t)ncu * per_cu;
  const dim3 grid((unsigned)(need < cap ? need : cap));
  return sink_launch(s, kernel, grid, dim3(block), io, tab);
}

// Output cache policy of int16 batches, linear and ragged.  Nontemporal
// stores stream the output past the caches; stores with the default policy
// leave it in the MI355X's 256 MB Infinity Cache, where the batch's next
// pass (the inverse after the forward, or the next transform in place) reads
// it instead of HBM.  That pays while the batch fits.  Measured, forward +
// inverse in place, MFMA16, per round trip:
//
//   polys        nontemporal   kept
//   32,768          17.7 us    20.1 us
//   65,536          34.5       29.8
//   262,144        106.3       92.6
//   524,288        201.9      166.1
//   589,824        219.9      241.0
//   2^20           361.1      382.2
//
// and, each pass timed alone after staging the batch with a device copy,
// 106.9 against 98.2 us at 262,144 polys.  MFMA32 gains from 98,304 polys
// in place (45.3 against 42.6 us) but loses 1-2.5 us per round trip below
// 131,072 polys when each pass follows a copy (65,536: 35.7 against 38.2).
// Batches of 32 MiB (MFMA16) or 64 MiB (MFMA32) up to 256 MiB keep their
// output.  Ragged plans gain alike: 262,144 polys in 127 requests, 110.7
// against 94.6 us in place, 109.4 against 103.3 with each pass timed alone.
constexpr size_t kKeepMaxPolys = 524288;

inline bool keep_polys(size_t npoly, int shape)
{
  return npoly >= (shape == 16 ? 65536u : 131072u) && npoly <= kKeepMaxPolys;
}
inline bool keep_output(const LinearTiles &io, int shape) { return keep_polys(io.npoly, shape); }
// Ragged plans: their tiles' polys, counting every tile as full (requests
// are mostly several tiles long).
inline bool keep_output(const RaggedTiles &io, int shape)
{
  return keep_polys(io.ntiles * (size_t)shape, shape);
}
template <class Tiles>
inline bool keep_output(const Tiles &, int)
{
  return false;
}

template <int SHAPE, int FORM, bool INV, bool KEEP, class Tiles>
int launch_policy(int waves, int ncu, int wgs, const Tiles &io, const DeviceTables *tab,
                  const LaunchSink &s)
{
  if (SHAPE == 16) {
    // The swap pays when a wave has more than one tile: the grid (capped at
    // `wgs` times the resident workgroups) then has fewer wave slots than
    // tiles.  Both instances have the same residency.
    auto pick = [&](auto swapped, auto plain, int nw) {
      const int per_cu = resident_per_cu((const void *)swapped, 64 * nw) * (wgs > 0 ? wgs : 1);
      const bool swap = host_tile_count(io, 16) > (size_t)ncu * per_cu * nw;
      return swap ? launch_persistent(swapped, nw, 16, ncu, wgs, io, tab, s)
                  : launch_persistent(plain, nw, 16, ncu, wgs, io, tab, s);
    };
    // Packed input: one workgroup per CU, the waves taking tiles from the
    // workgroup's queue (k_mfma16, QUEUED).
    if constexpr (PackedBits<Tiles>::value != 0) {
      const size_t ntiles = host_tile_count(io, 16);
      const dim3 grid((unsigned)(ntiles < (size_t)ncu ? ntiles : (size_t)ncu));
      return sink_launch(s, k_mfma16<FORM, INV, kQueuedWaves, Tiles, true, true>, grid,
                         dim3(64 * kQueuedWaves), io, tab);
    } else {
      return waves == 2 ? pick(k_mfma16<FORM, INV, 2, Tiles, true, false, KEEP>,
                               k_mfma16<FORM, INV, 2, Tiles, false, false, KEEP>, 2)
                        : pick(k_mfma16<FORM, INV, 4, Tiles, true, false, KEEP>,
                               k_mfma16<FORM, INV, 4, Tiles, false, false, KEEP>, 4);
    }
  }
  return waves == 2
             ? launch_persistent(k_mfma32<FORM, INV, 2, Tiles, KEEP>, 2, 32, ncu, wgs, io, tab, s)
             : launch_persistent(k_mfma32<FORM, INV, 4, Tiles, KEEP>, 4, 32, ncu, wgs, io, tab, s);
}

template <int SHAPE, int FORM, bool INV, class Tiles>
int launch_one(int waves, int ncu, int wgs, const Tiles &io, const DeviceTables *tab,
               const LaunchSink &s)
{
  if constexpr (std::is_same<Tiles, LinearTiles>::value || std::is_same<Tiles, RaggedTiles>::value)
    if (keep_output(io, SHAPE))
      return launch_policy<SHAPE, FORM, INV, true>(waves, ncu, wgs, io, tab, s);
  return launch_policy<SHAPE, FORM, INV, false>(waves, ncu, wgs, io, tab, s);
}

template <int SHAPE, bool INV, class Tiles>
int launch_form(int form, int waves, int ncu, int wgs, const Tiles &io, const DeviceTables *tab,
                const LaunchSink &s)
{
  switch (form) {
    case kFormCanon: return launch_one<SHAPE, kFormCanon, INV>(waves, ncu, wgs, io, tab, s);
    case kFormCentred: return launch_one<SHAPE, kFormCentred, INV>(waves, ncu, wgs, io, tab, s);
    default: return -1;
  }
}

template <class Tiles>
int launch_any(int shape, bool inverse, int form, const Tiles &io, const DeviceTables *tab,
               const LaunchShape &ls, const LaunchSink &stream)
{
  if (ls.waves != 2 && ls.waves != 4)
    return -1;
  if (shape == 32)
    return inverse ? launch_form<32, true>(form, ls.waves, ls.ncu, ls.oversub, io, tab, stream)
                   : launch_form<32, false>(form, ls.waves, ls.ncu, ls.oversub, io, tab, stream);
  return inverse ? launch_form<16, true>(form, ls.waves, ls.ncu, ls.oversub, io, tab, stream)
                 : launch_form<16, false>(form, ls.waves, ls.ncu, ls.oversub, io, tab, stream);
}

}  // namespace

int launch_mfma(int shape, bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                const DeviceTables *tab, const LaunchShape &ls, hipStream_t stream)
{
  const LinearTiles io{in, out, npoly};
  return launch_any(shape, inverse, form, io, tab, ls, stream);
}

void prepare_ragged_launch(int shape, int waves)
{
  // The forms ragged plans use: canonical forward, centred inverse; both
  // output policies (keep_output).
  auto one = [&](const void *k) { (void)resident_per_cu(k, 64 * waves); };
  auto both = [&](auto keep) {
    constexpr bool K = decltype(keep)::value;
    if (shape == 32) {
      one(waves == 2 ? (const void *)k_mfma32<kFormCanon, false, 2, RaggedTiles, K>
                     : (const void *)k_mfma32<kFormCanon, false, 4, RaggedTiles, K>);
      one(waves == 2 ? (const void *)k_mfma32<kFormCentred, true, 2, RaggedTiles, K>
                     : (const void *)k_mfma32<kFormCentred, true, 4, RaggedTiles, K>);
    } else {
      one(waves == 2 ? (const void *)k_mfma16<kFormCanon, false, 2, RaggedTiles, true, false, K>
                     : (const void *)k_mfma16<kFormCanon, false, 4, RaggedTiles, true, false, K>);
      one(waves == 2 ? (const void *)k_mfma16<kFormCentred, true, 2, RaggedTiles, true, false, K>
                     : (const void *)k_mfma16<kFormCentred, true, 4, RaggedTiles, true, false, K>);
    }
  };
  both(std::false_type{});
  both(std::true_type{});
}

int launch_mfma_ragged(int shape, bool inverse, int form, const RaggedTiles &io,
                       const DeviceTables *tab, const LaunchShape &ls, const LaunchSink &stream)
{
  return launch_any(shape, inverse, form, io, tab, ls, stream);
}

int launch_mfma_packed(int shape, int bits, int form, int16_t *out, const uint8_t *in,
                       size_t npoly, const DeviceTables *tab, const LaunchShape &ls,
                       hipStream_t stream)
{
  if (ls.waves != 2 && ls.waves != 4)
    return -1;
  auto go = [&](auto io) {
    return shape == 32 ? launch_form<32, false>(form, ls.waves, ls.ncu, ls.oversub, io, tab, stream)
                       : launch_form<16, false>(form, ls.waves, ls.ncu, ls.oversub, io, tab, stream);
  };
  switch (bits) {
    case 2: return go(PackedTiles<2>{in, out, npoly});
    case 3: return go(PackedTiles<3>{in, out, npoly});
    case 4: return go(PackedTiles<4>{in, out, npoly});
    default: return -1;
  }
}

}  // namespace tk
