// Ragged transform without allocation: the work of tk_ragged_plan_create +
// tk_ragged_plan_execute in one call, for callers that cannot allocate.
//
// The plan's device memory -- its table of tiles -- lives here in a
// file-scope device array, `scratch`, of a fixed size, one copy per device.
// A call copies the caller's request table into it, expands the requests
// into the table of tiles on the device (k_expand_tiles) and runs the ragged
// persistent kernel over that table with the linear batches' launch shape.
// Nothing is allocated: the array's device address is looked up once per
// device by tk_init (noalloc_init), which also caches the kernels' occupancy.
#include <hip/hip_runtime.h>

#include "tk_launch.h"
#include "tk_state.h"
#include "tkyber.h"

namespace tk {
namespace {

constexpr size_t kMaxReq = TK_NOALLOC_MAX_REQUESTS;
constexpr size_t kMaxTiles = TK_NOALLOC_MAX_TILES;
constexpr int kTileP = 16;  // MFMA16 tiles
static_assert(sizeof(tk_request) == 16 && sizeof(TileRef) == 16, "16-byte entries");

// [kMaxReq requests | kMaxTiles tiles]: 128 KiB + 2 MiB per device.
// In host code `scratch` names a host-side stand-in, not device memory:
// only the address from hipGetSymbolAddress (DevState::noalloc_scratch) is
// passed to kernels and copies.
static __device__ __align__(256) uint8_t scratch[(kMaxReq + kMaxTiles) * 16];

constexpr int kExpandThreads = 1024;
constexpr int kExpandPerThread = 4;  // tiles written per thread
constexpr int kReqPerThread = (int)(kMaxReq / kExpandThreads);

// Requests -> table of tiles of P polys (request order, each request's last
// tile partial).  Every workgroup scans the whole request table into LDS
// (first tile of each request), then writes its 4096 tiles, finding each
// tile's request by binary search.  Tiles the scan does not cover (ntiles
// above the requests' count: a table changed after the host counted it) get
// an empty entry, which the transform kernel reads as zero polys.
__global__ __launch_bounds__(kExpandThreads) void k_expand_tiles(const tk_request *__restrict__ reqs,
                                                                uint32_t nreq, int P,
                                                                TileRef *__restrict__ tiles,
                                                                uint32_t ntiles)
{
  __shared__ uint32_t s_first[kMaxReq + 1];  // first tile of request r; s_first[nreq] = total
  __shared__ uint32_t s_part[kExpandThreads];
  const int t = threadIdx.x;
  // Thread t counts requests t*kReqPerThread .. +kReqPerThread-1.
  uint32_t cnt[kReqPerThread], sum = 0;
#pragma unroll
  for (int i = 0; i < kReqPerThread; i++) {
    const uint32_t r = (uint32_t)t * kReqPerThread + i;
    cnt[i] = r < nreq ? (reqs[r].npoly + (uint32_t)P - 1) / (uint32_t)P : 0u;
    sum += cnt[i];
  }
  // Inclusive scan of the per-thread sums (Hillis-Steele in LDS).
  s_part[t] = sum;
  __syncthreads();
  for (int off = 1; off < kExpandThreads; off <<= 1) {
    const uint32_t v = t >= off ? s_part[t - off] : 0u;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  uint32_t run = s_part[t] - sum;  // exclusive prefix of thread t
#pragma unroll
  for (int i = 0; i < kReqPerThread; i++) {
    const uint32_t r = (uint32_t)t * kReqPerThread + i;
    if (r < nreq)
      s_first[r] = run;
    run += cnt[i];
  }
  if (t == 0)
    s_first[nreq] = s_part[kExpandThreads - 1];
  __syncthreads();
  const uint32_t total = s_first[nreq];
#pragma unroll
  for (int k = 0; k < kExpandPerThread; k++) {
    const uint32_t g = (blockIdx.x * kExpandPerThread + k) * kExpandThreads + t;
    if (g >= ntiles)
      break;
    TileRef e{nullptr, 0u, 0u};
    if (g < total) {
      // last r with s_first[r] <= g (empty requests share their successor's
      // first tile, so the search lands on the request that holds g)
      uint32_t lo = 0, hi = nreq;  // s_first[lo] <= g < s_first[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_first[mid] <= g)
          lo = mid;
        else
          hi = mid;
      }
      const uint32_t i = g - s_first[lo];
      const tk_request q = reqs[lo];
      const uint32_t done = i * (uint32_t)P;
      if (done < q.npoly) {
        const uint32_t n = q.npoly - done;
        e = TileRef{q.base + (size_t)done * kN, n < (uint32_t)P ? n : (uint32_t)P, 0u};
      }
    }
    tiles[g] = e;
  }
}

}  // namespace

int noalloc_init(DevState *st)
{
  void *p = nullptr;
  int rc = record_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (rc)
    return rc;
  st->noalloc_scratch = p;
  prepare_ragged_launch(kTileP, 4);
  return TK_OK;
}

}  // namespace tk

using namespace tk;

extern "C" {

int tk_ragged_transform_noalloc(const tk_request *table, size_t nreq, int direction, void *stream)
{
  if (!table || nreq == 0 || (direction != 0 && direction != 1))
    return TK_EINVAL;
  if (nreq > kMaxReq)
    return TK_EINVAL;
  size_t ntiles = 0;
  for (size_t r = 0; r < nreq; r++) {
    const tk_request &q = table[r];
    if (q.k < 2 || q.k > 4 || q.kind > 1)
      return TK_EINVAL;
    if (q.npoly && !aligned16(q.base))
      return TK_EINVAL;
    ntiles += (q.npoly + (size_t)kTileP - 1) / kTileP;
  }
  if (ntiles == 0 || ntiles > kMaxTiles)
    return TK_EINVAL;
  DevState *st = nullptr;
  int rc = device_ready(&st);
  if (rc)
    return rc;
  hipStream_t s = (hipStream_t)stream;
  tk_request *dreq = static_cast<tk_request *>(st->noalloc_scratch);
  TileRef *dtiles = reinterpret_cast<TileRef *>(dreq + kMaxReq);
  if ((rc = record_hip(hipMemcpyAsync(dreq, table, nreq * sizeof(tk_request),
                                      hipMemcpyHostToDevice, s))))
    return rc;
  const unsigned grid = (unsigned)((ntiles + kExpandThreads * kExpandPerThread - 1) /
                                   (kExpandThreads * kExpandPerThread));
  hipLaunchKernelGGL(k_expand_tiles, dim3(grid), dim3(kExpandThreads), 0, s, dreq,
                     (uint32_t)nreq, kTileP, dtiles, (uint32_t)ntiles);
  if ((rc = after_launch()))
    return rc;
  const int form = direction ? kFormCentred : kFormCanon;
  if (launch_mfma_ragged(kTileP, direction != 0, form, RaggedTiles{dtiles, ntiles}, st->tables,
                         LaunchShape{st->ncu, 4, kLinearOversub}, s))
    return TK_ENOTSUP;
  return after_launch();
}

}  // extern "C"
