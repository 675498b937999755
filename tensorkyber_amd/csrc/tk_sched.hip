// Rank-aware scheduling of mixed fused requests: ML-KEM-512/768/1024 use
// module ranks k = 2/3/4 (ref/params.h:10-18), and a serving batch mixes
// them with the operations (matvec, keygen, encryption, decryption).  A plan
// takes a table of requests, each a contiguous batch of one (operation, rank)
// with its own buffers, and
//   * groups the instances by (operation, rank): one fused kernel per group,
//     whatever the number of requests, each instance found through a device
//     table of its eight buffer pointers;
//   * orders the groups by work, largest first, and builds them as parallel
//     branches of one hipGraph (one kernel node each, no dependencies), so
//     small groups fill the CUs the large ones leave idle at their tails.
// Executing the plan replays the graph: one launch call per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "tk_state.h"
#include "tkyber.h"

namespace {

struct Group {
  int mode = 0, k = 0;
  size_t first = 0, ninst = 0;  // slice of the plan's instance table
};

}  // namespace

struct tk_fused_plan {
  int device = -1;
  std::vector<Group> groups;
  size_t ninst = 0;
  tk::FusedInst *table = nullptr;  // device
  const tk::DeviceTables *tab = nullptr;
  int ncu = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// Polys moved per instance, the work measure used to order the groups.
size_t polys_per_inst(int mode, int k)
{
  switch (mode) {
    case TK_FUSED_MATVEC: return (size_t)(k * k + 2 * k);
    case TK_FUSED_KEYGEN: return (size_t)(k * k + 4 * k);
    case TK_FUSED_ENCRYPT: return (size_t)(k * k + 4 * k + 3);
    default: return (size_t)(2 * k + 2);
  }
}

tk::FusedIo request_io(const tk_fused_request &r)
{
  tk::FusedIo io{};
  io.out0 = r.out0;
  io.out1 = r.out1;
  io.mat = r.mat;
  io.mat2 = r.mat2;
  io.x = r.x;
  io.add0 = r.add0;
  io.add1 = r.add1;
  io.add2 = r.add2;
  io.ninst = r.ninst;
  return io;
}

// Instance i of request r as table pointers (the contiguous layout's strides).
tk::FusedInst instance(const tk_fused_request &r, size_t i)
{
  const size_t kN = tk::kN, k = (size_t)r.k;
  const bool dec = r.mode == TK_FUSED_DECRYPT, enc = r.mode == TK_FUSED_ENCRYPT;
  auto at = [&](const int16_t *p, size_t stride) -> const int16_t * {
    return p ? p + i * stride : nullptr;
  };
  tk::FusedInst f{};
  f.out0 = const_cast<int16_t *>(at(r.out0, dec ? kN : k * kN));
  f.out1 = const_cast<int16_t *>(at(r.out1, enc ? kN : k * kN));
  f.mat = at(r.mat, dec ? k * kN : k * k * kN);
  f.mat2 = at(r.mat2, k * kN);
  f.x = at(r.x, k * kN);
  f.add0 = at(r.add0, dec ? kN : k * kN);
  f.add1 = at(r.add1, kN);
  f.add2 = at(r.add2, kN);
  return f;
}

void plan_free(tk_fused_plan *p)
{
  if (p->exec)
    (void)hipGraphExecDestroy(p->exec);
  if (p->graph)
    (void)hipGraphDestroy(p->graph);
  if (p->table)
    (void)hipFree(p->table);
  delete p;
}

int launch_group(const tk_fused_plan *p, const Group &g, const tk::LaunchSink &s)
{
  tk::FusedIo io{};
  io.ninst = g.ninst;
  io.inst = p->table + g.first;
  const int rc = tk::launch_fused(g.k, g.mode, 0, io, p->tab, p->ncu, s);
  if (rc)
    return rc == TK_EDEVICE ? rc : TK_ENOTSUP;
  return tk::after_launch();
}

// The groups as parallel branches of one graph: one kernel node each, with
// no dependencies between them, built node by node (no stream capture, so
// other host threads' device work cannot interfere with it).
int build_graph(tk_fused_plan *p)
{
  int rc = tk::record_hip(hipGraphCreate(&p->graph, 0));
  for (size_t g = 0; g < p->groups.size() && rc == TK_OK; g++)
    rc = launch_group(p, p->groups[g], tk::LaunchSink(p->graph));
  if (rc == TK_OK)
    rc = tk::record_hip(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  return rc;
}

}  // namespace

extern "C" {

int tk_fused_plan_create(tk_fused_plan **out, const tk_fused_request *reqs, size_t nreq)
{
  if (!out || !reqs || nreq == 0)
    return TK_EINVAL;
  *out = nullptr;
  // Validate every request as the single calls do, and count instances per
  // (operation, rank).
  size_t count[4][5] = {};
  size_t total = 0;
  std::vector<tk::ReqRange> outs, ins;
  try {
    for (size_t r = 0; r < nreq; r++) {
      const tk_fused_request &q = reqs[r];
      if (q.mode < TK_FUSED_MATVEC || q.mode > TK_FUSED_DECRYPT || q.k < 2 || q.k > 4)
        return TK_EINVAL;
      if (q.ninst == 0)
        continue;
      int rc = tk::validate_fused(q.k, q.mode, request_io(q));
      if (rc)
        return rc;
      count[q.mode][q.k] += q.ninst;
      total += q.ninst;
      if (total > tk::kMaxPolys / 16)
        return TK_EINVAL;
      tk::fused_ranges(q.k, q.mode, request_io(q), r, outs, ins);
    }
  } catch (const std::bad_alloc &) {
    return TK_ENOMEM;
  }
  if (total == 0)
    return TK_EINVAL;
  // The groups run concurrently and one kernel serves many requests: an
  // output shared with another request's buffers would be a data race.
  if (tk::cross_request_overlap(outs, ins))
    return TK_EINVAL;
  tk::DevState *st = nullptr;
  int rc = tk::init_device(&st);
  if (rc)
    return rc;

  tk_fused_plan *p = new tk_fused_plan();
  p->device = st->device;
  p->tab = st->tables;
  p->ncu = st->ncu;
  p->ninst = total;
  for (int m = 0; m < 4; m++)
    for (int k = 2; k <= 4; k++)
      if (count[m][k]) {
        Group g;
        g.mode = m;
        g.k = k;
        g.ninst = count[m][k];
        p->groups.push_back(g);
      }
  std::stable_sort(p->groups.begin(), p->groups.end(), [](const Group &a, const Group &b) {
    return a.ninst * polys_per_inst(a.mode, a.k) > b.ninst * polys_per_inst(b.mode, b.k);
  });
  std::vector<tk::FusedInst> host;
  try {
    host.resize(total);
  } catch (const std::bad_alloc &) {
    plan_free(p);
    return TK_ENOMEM;
  }
  size_t first = 0;
  for (Group &g : p->groups) {
    g.first = first;
    size_t at = first;
    for (size_t r = 0; r < nreq; r++)
      if (reqs[r].mode == g.mode && reqs[r].k == g.k)
        for (size_t i = 0; i < reqs[r].ninst; i++)
          host[at++] = instance(reqs[r], i);
    first += g.ninst;
  }
  if (hipMalloc(&p->table, total * sizeof(tk::FusedInst)) != hipSuccess) {
    plan_free(p);
    return TK_ENOMEM;
  }
  if ((rc = tk::record_hip(
           hipMemcpy(p->table, host.data(), total * sizeof(tk::FusedInst), hipMemcpyHostToDevice))) ||
      (rc = build_graph(p))) {
    plan_free(p);
    return rc;
  }
  *out = p;
  return TK_OK;
}

int tk_fused_plan_execute(tk_fused_plan *plan, void *stream)
{
  if (!plan || !plan->exec)
    return TK_EINVAL;
  return tk::record_hip(hipGraphLaunch(plan->exec, (hipStream_t)stream));
}

int tk_fused_plan_launch(tk_fused_plan *plan, void *stream)
{
  if (!plan)
    return TK_EINVAL;
  for (const Group &g : plan->groups) {
    const int rc = launch_group(plan, g, (hipStream_t)stream);
    if (rc)
      return rc;
  }
  return TK_OK;
}

int tk_fused_plan_info(const tk_fused_plan *plan, int *ngroups, size_t *ninst_total)
{
  if (!plan)
    return TK_EINVAL;
  if (ngroups)
    *ngroups = (int)plan->groups.size();
  if (ninst_total)
    *ninst_total = plan->ninst;
  return TK_OK;
}

int tk_fused_plan_destroy(tk_fused_plan *plan)
{
  if (!plan)
    return TK_EINVAL;
  plan_free(plan);
  return TK_OK;
}

}  // extern "C"
