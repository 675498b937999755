// Host-side launchers of the kernels (one per .hip file); called by the C ABI
// in tk_api.hip after argument validation.  Each returns 0, -1 for an
// unsupported form, or TK_EDEVICE when a graph node could not be added;
// HIP launch errors are read by the caller.
#pragma once

#include <hip/hip_runtime.h>

#include "tk_internal.h"

namespace tk {

int record_hip(hipError_t e);

// Where a launcher puts its kernel: launched on a stream, or added to a
// graph under construction as a kernel node with no dependencies.  Plans
// build their graphs this way rather than by stream capture: a capture is a
// window in which what other host threads do on the device can invalidate
// it or be refused.
struct LaunchSink {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  LaunchSink(hipStream_t s) : stream(s) {}  // NOLINT: a stream is a sink
  explicit LaunchSink(hipGraph_t g) : graph(g) {}
};

template <class T>
struct NoDeduce {
  using type = T;
};

// kernel<<<grid, block>>>(args...) into the sink: 0, or TK_EDEVICE if the
// graph node could not be added (launch errors are read by the caller).
template <class... P>
int sink_launch(const LaunchSink &sk, void (*kernel)(P...), dim3 grid, dim3 block,
                typename NoDeduce<P>::type... args)
{
  if (!sk.graph) {
    hipLaunchKernelGGL(kernel, grid, block, 0, sk.stream, args...);
    return 0;
  }
  void *params[] = {static_cast<void *>(&args)...};
  hipKernelNodeParams kp{};
  kp.func = reinterpret_cast<void *>(kernel);
  kp.gridDim = grid;
  kp.blockDim = block;
  kp.sharedMemBytes = 0;
  kp.kernelParams = params;
  kp.extra = nullptr;
  hipGraphNode_t node = nullptr;
  return record_hip(hipGraphAddKernelNode(&node, sk.graph, nullptr, 0, &kp));
}

// A contiguous batch: poly p at in + p*256 (read) and out + p*256 (written).
struct LinearTiles {
  const int16_t *in;
  int16_t *out;
  size_t npoly;
};

// A ragged batch of requests, transformed in place, as a table of tiles:
// each request is cut into tiles of the kernel's tile height P (16 or 32
// polys), the last one partial, and tiles[t] holds tile t's first poly and
// its number of polys (1..P).  A wave reads its tile's entry with one scalar
// load, issued a tile ahead, and stages the tile through a buffer descriptor
// over its polys as a linear batch does.  16 bytes per tile (1 byte per poly
// at P = 16) where a table of per-poly addresses took 8 per poly.
struct TileRef {
  int16_t *base;
  uint32_t npoly;
  uint32_t reserved;
};
static_assert(sizeof(TileRef) == 16, "one 16-byte scalar load per tile");
struct RaggedTiles {
  const TileRef *tiles;
  size_t ntiles;
};

// A contiguous batch of small-coefficient polys in the packed BITS-bit form
// (see packed_poly_bytes): poly p at in + p * packed bytes, its forward
// transform written to out + p*256.
template <int BITS>
struct PackedTiles {
  const uint8_t *in;
  int16_t *out;
  size_t npoly;
};

// Grid shape of the tile kernels: waves per workgroup, and the grid as a
// multiple of the workgroups resident at once (from the occupancy
// calculator).  Each wave strides over tiles; a factor above 1 lets the
// dispatcher hand late workgroups to whichever CU frees up first, which
// balances the load.  Linear batches use 8 for 16-poly tiles (2^20 polys:
// 181-183 us per pass, against 187-188 at 4 and 189-191 at 16) and 12 for
// 32-poly tiles (180-183 / 184-188 us forward / inverse, against 188-190 /
// 190-192 at 8; the inverse slows again from 16).  Packed batches load few
// bytes per tile: 32-poly tiles use 1, with workgroups of 2 waves for 2-bit
// input and 4 for 3 and 4 bits; 16-poly tiles run one 16-wave workgroup per
// CU whose waves take tiles from a queue in LDS (k_mfma16, QUEUED): 2^20
// polys take 104-105 us at 2 bits and 108 us at 3, where the 1x grid of 2-
// or 4-wave workgroups striding over the tiles took 126 and 135 us (and 4x
// 141-143 us at 2 bits) with nontemporal loads.  Ragged plans let the
// autotuner choose, starting from the linear shape.
constexpr int kLinearOversub = 8;
constexpr int kLinearOversub32 = 12;
constexpr int kPackedOversub = 1;
constexpr int packed_waves(int bits) { return bits == 2 ? 2 : 4; }
struct LaunchShape {
  int ncu;
  int waves;
  int oversub;
};

// Workgroups of `block_threads` that fit on one CU at once (cached).
int resident_per_cu(const void *kernel, int block_threads);

int launch_butterfly(bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                     const DeviceTables *tab, int ncu, hipStream_t stream);

int launch_mfma(int shape, bool inverse, int form, int16_t *out, const int16_t *in, size_t npoly,
                const DeviceTables *tab, const LaunchShape &ls, hipStream_t stream);

// Caches the occupancy of the ragged kernels of one shape and workgroup
// size, so that their later launches allocate nothing.
void prepare_ragged_launch(int shape, int waves);

// io.tiles must have been cut for shape's tile height (16 or 32 polys).
int launch_mfma_ragged(int shape, bool inverse, int form, const RaggedTiles &io,
                       const DeviceTables *tab, const LaunchShape &ls, const LaunchSink &sink);

// Forward transform of packed small-coefficient polys (bits 2, 3 or 4).
int launch_mfma_packed(int shape, int bits, int form, int16_t *out, const uint8_t *in,
                       size_t npoly, const DeviceTables *tab, const LaunchShape &ls,
                       hipStream_t stream);

// Packs int16 polys into the BITS-bit form (bits 2, 3 or 4).
int launch_pack_small(int bits, uint8_t *out, const int16_t *in, size_t npoly, int ncu,
                      hipStream_t stream);

// Modes of the fused module kernel and the buffers each uses (shapes per
// instance; k = module rank; all polys int16[256] unless packed):
//   kFusedMatvec:  out0 = t [k] <- invntt(mat [k][k] . ntt(x [k]))
//   kFusedKeygen:  out0 = t_hat [k], out1 = s_hat [k] <- mat = A_hat [k][k],
//                  x = s [k], add0 = e [k]
//   kFusedEncrypt: out0 = u [k], out1 = v [1] <- mat = At_hat [k][k],
//                  mat2 = t_hat [k], x = r [k], add0 = e1 [k], add1 = e2 [1],
//                  add2 = m [1] or null
//   kFusedDecrypt: out0 = w [1] <- mat = s_hat [k], x = u [k], add0 = v [1]
enum FusedMode { kFusedMatvec = 0, kFusedKeygen = 1, kFusedEncrypt = 2, kFusedDecrypt = 3 };

// One instance's buffers for a table-driven launch, in FusedIo's roles: each
// points at that instance's first poly of the role (the role's shape as
// above); add2 may be null.  Eight pointers, read as an array by role.
struct FusedInst {
  int16_t *out0;
  int16_t *out1;
  const int16_t *mat;
  const int16_t *mat2;
  const int16_t *x;
  const int16_t *add0;
  const int16_t *add1;
  const int16_t *add2;
};
static_assert(sizeof(FusedInst) == 8 * sizeof(void *), "eight pointers");

struct FusedIo {
  int16_t *out0;
  int16_t *out1;
  const int16_t *mat;
  const int16_t *mat2;
  const void *x;  // int16 polys, or packed (matvec only)
  const int16_t *add0;
  const int16_t *add1;
  const int16_t *add2;
  size_t ninst;
  // Non-null: instance g's buffers are inst[g] (int16 x only) and the
  // pointers above are unused.
  const FusedInst *inst;
};

// s_bits = 0: x is int16 polys; 2, 3 or 4: x is packed (kFusedMatvec only).
int launch_fused(int k, int mode, int s_bits, const FusedIo &io, const DeviceTables *tab, int ncu,
                 const LaunchSink &sink);

int launch_basemul(int form, int16_t *r, const int16_t *a, const int16_t *b, size_t npoly,
                   const DeviceTables *tab, int ncu, hipStream_t stream);

}  // namespace tk
