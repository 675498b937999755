// Ragged batches (many uneven ML-KEM-512/768/1024 vectors and matrices in
// one launch) and the launch-shape autotuner.
//
// A plan flattens the request table into a global poly numbering, uploads
// the device address of every poly once, picks the MFMA tile shape / waves
// per workgroup / grid with the autotuner, and records the persistent kernel
// into a hipGraph.  Executing the plan replays
// the graph: one launch whatever the number of requests, no host work.
#include <hip/hip_runtime.h>
#include <string.h>

#include <map>
#include <new>
#include <mutex>
#include <tuple>
#include <vector>

#include "tk_launch.h"
#include "tk_state.h"
#include "tkyber.h"

struct tk_ragged_plan {
  int device = -1;
  int direction = 0;
  int shape = 16;
  int waves = 4;
  int ncu = 0;
  int oversub = 0;
  size_t npoly = 0;
  size_t nreq = 0;
  void *dmem = nullptr;
  tk::RaggedTiles tiles{};
  const tk::DeviceTables *tab = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace tk {
namespace {

struct TuneKey {
  int device, direction, bucket;
  bool operator<(const TuneKey &o) const
  {
    return std::tie(device, direction, bucket) < std::tie(o.device, o.direction, o.bucket);
  }
};
struct TuneVal {
  int method, waves, oversub;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

int bucket_of(size_t npoly)
{
  int b = 0;
  while (b < 40 && ((size_t)1 << (b + 1)) <= npoly)
    b++;
  return b;
}

int time_config(DevState *st, int shape, int waves, int wgs, bool inverse, const RaggedTiles &io,
                hipStream_t s, float *ms_out)
{
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess)
    return TK_EDEVICE;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return TK_EDEVICE;
  }
  const int form = inverse ? kFormCentred : kFormCanon;
  const LaunchShape ls{st->ncu, waves, wgs};
  // Two warm-up launches, then the best of three groups of four launches
  // back to back (per-launch time), which rides out clock and queue noise.
  constexpr int kGroups = 3, kPerGroup = 4;
  float best = 1e30f;
  int rc = TK_OK;
  for (int w = 0; w < 2; w++)
    launch_mfma_ragged(shape, inverse, form, io, st->tables, ls, s);
  for (int grp = 0; grp < kGroups && rc == TK_OK; grp++) {
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < kPerGroup; i++)
      launch_mfma_ragged(shape, inverse, form, io, st->tables, ls, s);
    (void)hipEventRecord(e1, s);
    if ((rc = record_hip(hipEventSynchronize(e1))) != TK_OK)
      break;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms / kPerGroup < best)
      best = ms / kPerGroup;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_out = best;
  return rc;
}

int autotune(DevState *st, size_t npoly, int direction, TuneVal *out, hipStream_t s)
{
  const TuneKey key{st->device, direction, bucket_of(npoly)};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      *out = it->second;
      return TK_OK;
    }
  }
  // Time the plans' own kernels (per-poly address table) on scratch data of
  // the same size (capped at 2^20 polys): n contiguous polys, the table
  // padded with a scratch poly as a plan's is.  The kernels' cost does not
  // depend on the values.  Timing the linear-batch kernels instead ranked
  // the shapes differently from the plans they were chosen for.
  const size_t n = npoly < ((size_t)1 << 20) ? npoly : ((size_t)1 << 20);
  const size_t padded = (n + kRaggedPad - 1) / kRaggedPad * kRaggedPad;
  const size_t b_data = (n + 1) * kN * sizeof(int16_t), b_ptrs = padded * sizeof(int16_t *);
  void *mem = nullptr;
  if (hipMalloc(&mem, b_data + b_ptrs) != hipSuccess)
    return TK_ENOMEM;
  int16_t *buf = static_cast<int16_t *>(mem);
  int16_t **dptrs = reinterpret_cast<int16_t **>(static_cast<char *>(mem) + b_data);
  std::vector<int16_t *> ptrs;
  try {
    ptrs.resize(padded);
  } catch (...) {
    (void)hipFree(mem);
    return TK_ENOMEM;
  }
  for (size_t i = 0; i < padded; i++)
    ptrs[i] = buf + (i < n ? i : n) * kN;
  int rc = record_hip(hipMemsetAsync(buf, 0, b_data, s));
  if (rc == TK_OK)
    rc = record_hip(hipMemcpyAsync(dptrs, ptrs.data(), b_ptrs, hipMemcpyHostToDevice, s));
  const RaggedTiles io{dptrs, n};
  // Candidates: tile shape x waves per workgroup x grid oversubscription
  // (grid = factor x workgroups resident at once).
  TuneVal best{TK_METHOD_MFMA16, 4, kDefaultOversub};
  float best_ms = 1e30f, default_ms = 1e30f;
  const int shapes[2] = {16, 32};
  const int wavess[2] = {2, 4};
  const int wgss[4] = {1, 2, 4, 8};
  for (int si = 0; si < 2 && rc == TK_OK; si++)
    for (int wi = 0; wi < 2 && rc == TK_OK; wi++)
      for (int gi = 0; gi < 4 && rc == TK_OK; gi++) {
        const int shape = shapes[si], waves = wavess[wi], wgs = wgss[gi];
        float ms = 0;
        rc = time_config(st, shape, waves, wgs, direction != 0, io, s, &ms);
        if (rc == TK_OK && shape == 16 && waves == 4 && wgs == kDefaultOversub)
          default_ms = ms;
        if (rc == TK_OK && ms < best_ms) {
          best_ms = ms;
          best = TuneVal{shape == 32 ? TK_METHOD_MFMA32 : TK_METHOD_MFMA16, waves, wgs};
        }
      }
  // Keep the default shape unless another is clearly faster: differences
  // inside the timing noise would otherwise make the choice vary by run.
  if (default_ms <= best_ms * 1.01f)
    best = TuneVal{TK_METHOD_MFMA16, 4, kDefaultOversub};
  (void)hipStreamSynchronize(s);
  (void)hipFree(mem);
  if (rc != TK_OK)
    return rc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = best;
  *out = best;
  return TK_OK;
}

void plan_free(tk_ragged_plan *p)
{
  if (p->exec)
    (void)hipGraphExecDestroy(p->exec);
  if (p->graph)
    (void)hipGraphDestroy(p->graph);
  if (p->dmem)
    (void)hipFree(p->dmem);
  delete p;
}

int plan_launch(tk_ragged_plan *p, hipStream_t s)
{
  const int form = p->direction ? kFormCentred : kFormCanon;
  if (launch_mfma_ragged(p->shape, p->direction != 0, form, p->tiles, p->tab,
                         LaunchShape{p->ncu, p->waves, p->oversub}, s))
    return TK_ENOTSUP;
  return after_launch();
}

}  // namespace
}  // namespace tk

using namespace tk;

extern "C" {

int tk_autotune(size_t npoly, int direction, int *best_method, int *best_waves_per_wg,
                int *best_oversub, void *stream)
{
  if (npoly == 0 || (direction != 0 && direction != 1) || npoly > kMaxPolys)
    return TK_EINVAL;
  DevState *st = nullptr;
  int rc = init_device(&st);
  if (rc)
    return rc;
  TuneVal v{};
  if ((rc = autotune(st, npoly, direction, &v, (hipStream_t)stream)))
    return rc;
  if (best_method)
    *best_method = v.method;
  if (best_waves_per_wg)
    *best_waves_per_wg = v.waves;
  if (best_oversub)
    *best_oversub = v.oversub;
  return TK_OK;
}

int tk_ragged_plan_create(tk_ragged_plan **plan, const tk_request *table, size_t nreq,
                          int direction, int method, void *stream)
{
  if (!plan || !table || nreq == 0 || nreq > 0xffffffffu || (direction != 0 && direction != 1))
    return TK_EINVAL;
  if (method != TK_METHOD_AUTO && method != TK_METHOD_MFMA && method != TK_METHOD_MFMA16 &&
      method != TK_METHOD_MFMA32)
    return TK_EINVAL;
  *plan = nullptr;
  size_t total = 0;
  for (size_t r = 0; r < nreq; r++) {
    const tk_request &q = table[r];
    if (q.k < 2 || q.k > 4 || q.kind > 1)
      return TK_EINVAL;
    if (q.npoly && !aligned16(q.base))
      return TK_EINVAL;
    total += q.npoly;
    if (total > kMaxPolys)
      return TK_EINVAL;
  }
  if (total == 0)
    return TK_EINVAL;
  // Every request is transformed in place by one kernel: a poly in two
  // requests would be transformed twice, concurrently.
  {
    std::vector<ReqRange> outs;
    try {
      outs.reserve(nreq);
    } catch (const std::bad_alloc &) {
      return TK_ENOMEM;
    }
    for (size_t r = 0; r < nreq; r++)
      if (table[r].npoly)
        outs.push_back(ReqRange{(uintptr_t)table[r].base,
                                (uintptr_t)table[r].base + (size_t)table[r].npoly * kN * 2, r});
    if (cross_request_overlap(outs, {}))
      return TK_EINVAL;
  }
  DevState *st = nullptr;
  int rc = init_device(&st);
  if (rc)
    return rc;
  hipStream_t s = (hipStream_t)stream;

  TuneVal tv{TK_METHOD_MFMA16, 4, kDefaultOversub};
  if (method == TK_METHOD_AUTO) {
    if ((rc = autotune(st, total, direction, &tv, s)))
      return rc;
  } else {
    tv.method = method == TK_METHOD_MFMA32 ? TK_METHOD_MFMA32 : TK_METHOD_MFMA16;
    tv.waves = 4;
    tv.oversub = kDefaultOversub;
  }
  const int shape = tv.method == TK_METHOD_MFMA32 ? 32 : 16;

  // Host image of the address table: poly g of the flattened numbering,
  // padded to a multiple of kRaggedPad entries that point at the plan's
  // scratch poly (after the table in the same allocation).
  const size_t padded = (total + kRaggedPad - 1) / kRaggedPad * kRaggedPad;
  const size_t b_ptrs = padded * sizeof(int16_t *);
  const size_t b_alloc = b_ptrs + kN * sizeof(int16_t);
  std::vector<int16_t *> ptrs;
  try {
    ptrs.resize(padded);
  } catch (const std::bad_alloc &) {
    return TK_ENOMEM;
  }
  size_t g = 0;
  for (size_t r = 0; r < nreq; r++)
    for (uint32_t i = 0; i < table[r].npoly; i++)
      ptrs[g++] = table[r].base + (size_t)i * kN;

  tk_ragged_plan *p = new tk_ragged_plan();
  p->device = st->device;
  p->direction = direction;
  p->shape = shape;
  p->waves = tv.waves;
  p->ncu = st->ncu;
  p->oversub = tv.oversub;
  p->npoly = total;
  p->nreq = nreq;
  p->tab = st->tables;
  if (hipMalloc(&p->dmem, b_alloc) != hipSuccess) {
    plan_free(p);
    return TK_ENOMEM;
  }
  int16_t *scratch = reinterpret_cast<int16_t *>(static_cast<char *>(p->dmem) + b_ptrs);
  for (size_t i = total; i < padded; i++)
    ptrs[i] = scratch;
  if ((rc = record_hip(hipMemcpy(p->dmem, ptrs.data(), b_ptrs, hipMemcpyHostToDevice)))) {
    plan_free(p);
    return rc;
  }
  p->tiles = RaggedTiles{(int16_t *const *)p->dmem, total};

  // Record the launch into a graph on a private capture stream.
  hipStream_t cap = nullptr;
  if ((rc = record_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking)))) {
    plan_free(p);
    return rc;
  }
  rc = record_hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  if (rc == TK_OK) {
    const int lrc = plan_launch(p, cap);
    const int erc = record_hip(hipStreamEndCapture(cap, &p->graph));
    rc = lrc ? lrc : erc;
  }
  if (rc == TK_OK)
    rc = record_hip(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  (void)hipStreamDestroy(cap);
  if (rc) {
    plan_free(p);
    return rc;
  }
  *plan = p;
  return TK_OK;
}

int tk_ragged_plan_execute(tk_ragged_plan *plan, void *stream)
{
  if (!plan || !plan->exec)
    return TK_EINVAL;
  return record_hip(hipGraphLaunch(plan->exec, (hipStream_t)stream));
}

int tk_ragged_plan_launch(tk_ragged_plan *plan, void *stream)
{
  if (!plan)
    return TK_EINVAL;
  return plan_launch(plan, (hipStream_t)stream);
}

int tk_ragged_plan_info(const tk_ragged_plan *plan, int *method, int *polys_per_wave, int *grid,
                        size_t *npoly_total)
{
  if (!plan)
    return TK_EINVAL;
  if (method)
    *method = plan->shape == 32 ? TK_METHOD_MFMA32 : TK_METHOD_MFMA16;
  if (polys_per_wave)
    *polys_per_wave = plan->shape;
  if (grid)
    *grid = plan->oversub;
  if (npoly_total)
    *npoly_total = plan->npoly;
  return TK_OK;
}

int tk_ragged_plan_destroy(tk_ragged_plan *plan)
{
  if (!plan)
    return TK_EINVAL;
  plan_free(plan);
  return TK_OK;
}

}  // extern "C"
