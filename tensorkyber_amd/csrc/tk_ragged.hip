// Ragged batches (many uneven ML-KEM-512/768/1024 vectors and matrices in
// one launch) and the launch-shape autotuner.
//
// A plan cuts every request into tiles of the kernel's tile height (16 or
// 32 polys, the last tile of a request partial) and uploads the table of
// tiles once: 16 bytes per tile, {first poly, number of polys}.  The
// persistent MFMA kernel walks that table (one scalar load per tile, issued
// a tile ahead) and stages each tile through a buffer descriptor over its
// polys exactly as it stages a linear batch.  The autotuner picks the tile
// shape, waves per workgroup and grid from timings of the plan's own
// requests.  Executing the plan is one dispatch of that kernel whatever the
// number of requests, with no host work; the launch is also the one kernel
// node of a hipGraph, which callers can embed in their own graphs.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
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
  hipGraph_t graph = nullptr;  // the recorded launch, for composition
};

namespace tk {
namespace {

struct TuneKey {
  int device, direction, bucket, req_bucket;
  bool operator<(const TuneKey &o) const
  {
    return std::tie(device, direction, bucket, req_bucket) <
           std::tie(o.device, o.direction, o.bucket, o.req_bucket);
  }
};
struct TuneVal {
  int method, waves, oversub;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

int octave(size_t n)
{
  int b = 0;
  while (b < 40 && ((size_t)1 << (b + 1)) <= n)
    b++;
  return b;
}

// The shape linear batches use: MFMA16 tiles, 4 waves per workgroup, a grid
// of 8x the resident workgroups.  Another shape must beat it clearly: by
// 5%.  On MI355X the tuner's groups have timed MFMA32 at 12x 2-3% ahead of
// it on config 4, while the plans themselves, timed interleaved, ran within
// 0.4% of each other (0.3713 against 0.3701 ms per round trip).
constexpr TuneVal kBaseline{TK_METHOD_MFMA16, 4, kLinearOversub};
constexpr float kSwitchMargin = 0.95f;  // switch only for >= 5% less time

// Cuts requests (base, npoly) into tiles of P polys.
void cut_tiles(const tk_request *table, size_t nreq, int P, std::vector<TileRef> &out)
{
  out.clear();
  for (size_t r = 0; r < nreq; r++)
    for (uint32_t i = 0; i < table[r].npoly; i += (uint32_t)P) {
      const uint32_t n = table[r].npoly - i;
      out.push_back(TileRef{table[r].base + (size_t)i * kN, n < (uint32_t)P ? n : (uint32_t)P, 0});
    }
}

size_t count_tiles(const tk_request *table, size_t nreq, int P)
{
  size_t n = 0;
  for (size_t r = 0; r < nreq; r++)
    n += (table[r].npoly + (size_t)P - 1) / (size_t)P;
  return n;
}

struct Candidate {
  int shape, waves, oversub;
  float ms;                // median over the rounds
  std::vector<float> runs;  // per round
};

// One group of launches back to back on s: two untimed ones, so that the
// timed ones start on a busy GPU as in a stream of work, then kPerGroup
// timed; per-launch time.
int time_group(DevState *st, const Candidate &c, bool inverse, const RaggedTiles &io,
               hipStream_t s, hipEvent_t e0, hipEvent_t e1, float *ms)
{
  constexpr int kLead = 2, kPerGroup = 6;
  const int form = inverse ? kFormCentred : kFormCanon;
  const LaunchShape ls{st->ncu, c.waves, c.oversub};
  for (int i = 0; i < kLead; i++)
    launch_mfma_ragged(c.shape, inverse, form, io, st->tables, ls, s);
  (void)hipEventRecord(e0, s);
  for (int i = 0; i < kPerGroup; i++)
    launch_mfma_ragged(c.shape, inverse, form, io, st->tables, ls, s);
  (void)hipEventRecord(e1, s);
  int rc = record_hip(hipEventSynchronize(e1));
  if (rc == TK_OK)
    rc = after_launch();
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  *ms = t / kPerGroup;
  return rc;
}

// Times every candidate on requests laid out as the plan's are and returns
// the fastest, or the baseline unless another beats it by the margin.
//
// The requests are rebased onto a scratch pool: with their own gaps when the
// pool spans at most 1.25x their bytes (the kernels' cost does not depend on
// the values), else back to back.  At most 2^20 polys of requests are timed
// (a prefix of the table).  The candidates are timed in interleaved rounds
// and scored by their median round, so a clock ramp or a noisy neighbour
// affects them alike and one lucky group does not decide.
int autotune_requests(DevState *st, const tk_request *table, size_t nreq, int direction,
                      TuneVal *out, hipStream_t s)
{
  constexpr size_t kCap = (size_t)1 << 20;
  size_t total = 0, used = 0;
  for (size_t r = 0; r < nreq; r++)
    total += table[r].npoly;
  const TuneKey key{st->device, direction, octave(total), octave(total / nreq)};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      *out = it->second;
      return TK_OK;
    }
  }
  std::vector<tk_request> reqs;
  std::vector<TileRef> t16, t32;
  try {
    size_t taken = 0;
    for (size_t r = 0; r < nreq && taken < kCap; r++) {
      if (!table[r].npoly)
        continue;
      tk_request q = table[r];
      if (q.npoly > kCap - taken)
        q.npoly = (uint32_t)(kCap - taken);
      taken += q.npoly;
      reqs.push_back(q);
    }
    used = taken;
  } catch (const std::bad_alloc &) {
    return TK_ENOMEM;
  }
  uintptr_t lo = UINTPTR_MAX, hi = 0;
  for (const tk_request &q : reqs) {
    lo = (uintptr_t)q.base < lo ? (uintptr_t)q.base : lo;
    const uintptr_t e = (uintptr_t)q.base + (size_t)q.npoly * kN * 2;
    hi = e > hi ? e : hi;
  }
  const size_t bytes = used * kN * 2;
  const bool keep_gaps = hi - lo <= bytes + bytes / 4;
  const size_t pool_bytes = keep_gaps ? hi - lo : bytes;
  void *pool = nullptr;
  if (hipMalloc(&pool, pool_bytes) != hipSuccess)
    return TK_ENOMEM;
  size_t at = 0;
  for (tk_request &q : reqs) {
    const size_t off = keep_gaps ? (uintptr_t)q.base - lo : at;
    q.base = reinterpret_cast<int16_t *>(static_cast<char *>(pool) + off);
    at += (size_t)q.npoly * kN * 2;
  }
  void *dt = nullptr;
  int rc = TK_OK;
  try {
    cut_tiles(reqs.data(), reqs.size(), 16, t16);
    cut_tiles(reqs.data(), reqs.size(), 32, t32);
  } catch (const std::bad_alloc &) {
    rc = TK_ENOMEM;
  }
  const size_t b16 = t16.size() * sizeof(TileRef), b32 = t32.size() * sizeof(TileRef);
  if (rc == TK_OK && hipMalloc(&dt, b16 + b32) != hipSuccess)
    rc = TK_ENOMEM;
  if (rc == TK_OK)
    rc = record_hip(hipMemsetAsync(pool, 0, pool_bytes, s));
  if (rc == TK_OK)
    rc = record_hip(hipMemcpyAsync(dt, t16.data(), b16, hipMemcpyHostToDevice, s));
  if (rc == TK_OK)
    rc = record_hip(hipMemcpyAsync(static_cast<char *>(dt) + b16, t32.data(), b32,
                                   hipMemcpyHostToDevice, s));
  const RaggedTiles io16{static_cast<const TileRef *>(dt), t16.size()};
  const RaggedTiles io32{reinterpret_cast<const TileRef *>(static_cast<char *>(dt) + b16),
                         t32.size()};

  std::vector<Candidate> cand;
  for (int shape : {16, 32})
    for (int waves : {2, 4})
      for (int oversub : {4, 8, 12})
        cand.push_back(Candidate{shape, waves, oversub, 1e30f, {}});
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == TK_OK)
    rc = record_hip(hipEventCreate(&e0));
  if (rc == TK_OK)
    rc = record_hip(hipEventCreate(&e1));
  const bool inverse = direction != 0;
  // Warm-up (clocks, code objects), then interleaved rounds.
  constexpr int kRounds = 5;
  for (int round = -1; round < kRounds && rc == TK_OK; round++)
    for (Candidate &c : cand) {
      float ms = 0;
      rc = time_group(st, c, inverse, c.shape == 32 ? io32 : io16, s, e0, e1, &ms);
      if (rc != TK_OK)
        break;
      if (round >= 0)
        c.runs.push_back(ms);
    }
  for (Candidate &c : cand)
    if (!c.runs.empty()) {
      std::sort(c.runs.begin(), c.runs.end());
      c.ms = c.runs[c.runs.size() / 2];
    }
  if (e0)
    (void)hipEventDestroy(e0);
  if (e1)
    (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  if (dt)
    (void)hipFree(dt);
  (void)hipFree(pool);
  if (rc != TK_OK)
    return rc;
  const Candidate *best = &cand[0], *base = nullptr;
  for (const Candidate &c : cand) {
    if (c.ms < best->ms)
      best = &c;
    if (c.shape == 16 && c.waves == kBaseline.waves && c.oversub == kBaseline.oversub)
      base = &c;
  }
  if (getenv("TK_DEBUG"))
    for (const Candidate &c : cand)
      fprintf(stderr, "tkyber autotune dir %d: shape %d waves %d oversub %2d: %.4f ms\n",
              direction, c.shape, c.waves, c.oversub, c.ms);
  TuneVal v = kBaseline;
  if (base && best != base && best->ms < base->ms * kSwitchMargin)
    v = TuneVal{best->shape == 32 ? TK_METHOD_MFMA32 : TK_METHOD_MFMA16, best->waves,
                best->oversub};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = v;
  *out = v;
  return TK_OK;
}

void plan_free(tk_ragged_plan *p)
{
  if (p->graph)
    (void)hipGraphDestroy(p->graph);
  if (p->dmem)
    (void)hipFree(p->dmem);
  delete p;
}

// The plan's kernel on a stream, or as the node of its graph.
int plan_launch(tk_ragged_plan *p, const LaunchSink &s)
{
  const int form = p->direction ? kFormCentred : kFormCanon;
  const int rc = launch_mfma_ragged(p->shape, p->direction != 0, form, p->tiles, p->tab,
                                    LaunchShape{p->ncu, p->waves, p->oversub}, s);
  if (rc)
    return rc == TK_EDEVICE ? rc : TK_ENOTSUP;
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
  // One contiguous request of npoly polys (at most 2^20 are timed).
  const size_t n = npoly < ((size_t)1 << 20) ? npoly : ((size_t)1 << 20);
  tk_request one{reinterpret_cast<int16_t *>((uintptr_t)256), (uint32_t)n, 2, 0};
  TuneVal v{};
  if ((rc = autotune_requests(st, &one, 1, direction, &v, (hipStream_t)stream)))
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

  TuneVal tv = kBaseline;
  if (method == TK_METHOD_AUTO) {
    if ((rc = autotune_requests(st, table, nreq, direction, &tv, s)))
      return rc;
  } else if (method == TK_METHOD_MFMA32) {
    tv = TuneVal{TK_METHOD_MFMA32, 4, kLinearOversub32};
  }
  const int shape = tv.method == TK_METHOD_MFMA32 ? 32 : 16;

  std::vector<TileRef> tiles;
  try {
    tiles.reserve(count_tiles(table, nreq, shape));
    cut_tiles(table, nreq, shape, tiles);
  } catch (const std::bad_alloc &) {
    return TK_ENOMEM;
  }
  const size_t b_tiles = tiles.size() * sizeof(TileRef);

  tk_ragged_plan *p = new (std::nothrow) tk_ragged_plan();
  if (!p)
    return TK_ENOMEM;
  p->device = st->device;
  p->direction = direction;
  p->shape = shape;
  p->waves = tv.waves;
  p->ncu = st->ncu;
  p->oversub = tv.oversub;
  p->npoly = total;
  p->nreq = nreq;
  p->tab = st->tables;
  if (hipMalloc(&p->dmem, b_tiles) != hipSuccess) {
    plan_free(p);
    return TK_ENOMEM;
  }
  if ((rc = record_hip(hipMemcpy(p->dmem, tiles.data(), b_tiles, hipMemcpyHostToDevice)))) {
    plan_free(p);
    return rc;
  }
  p->tiles = RaggedTiles{static_cast<const TileRef *>(p->dmem), tiles.size()};

  // The launch as the one kernel node of the plan's graph, built node by
  // node (no stream capture, so other host threads' device work cannot
  // interfere with it).
  rc = record_hip(hipGraphCreate(&p->graph, 0));
  if (rc == TK_OK)
    rc = plan_launch(p, LaunchSink(p->graph));
  if (rc) {
    plan_free(p);
    return rc;
  }
  *plan = p;
  return TK_OK;
}

// One dispatch of the recorded persistent kernel.  A replay of the plan's
// one-node graph would cost more: each graph launch adds 4-10 us of device
// time on this runtime (an empty kernel: 2.4 us direct, 11.5 us as a graph;
// a 190 us copy: +4 us; eight copies in one graph: +0 per copy), about 2.5%
// of a config-4 pass.  The graph is for callers that compose plans into
// their own graphs (tk_ragged_plan_graph), where that cost is paid once per
// graph.
int tk_ragged_plan_execute(tk_ragged_plan *plan, void *stream)
{
  if (!plan)
    return TK_EINVAL;
  return plan_launch(plan, (hipStream_t)stream);
}

int tk_ragged_plan_launch(tk_ragged_plan *plan, void *stream)
{
  return tk_ragged_plan_execute(plan, stream);
}

int tk_ragged_plan_graph(tk_ragged_plan *plan, void **graph)
{
  if (!plan || !graph || !plan->graph)
    return TK_EINVAL;
  *graph = plan->graph;
  return TK_OK;
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
