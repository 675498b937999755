// C ABI (include/tkyber.h): argument validation, per-device constant tables,
// method/form resolution and dispatch to the kernel launchers.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <utility>

#include "tk_internal.h"
#include "tk_launch.h"
#include "tk_state.h"
#include "tkyber.h"

namespace tk {
namespace {

std::mutex g_mu;
DevState g_dev[kMaxDevices];
HostMatrices g_hm;
DeviceTables g_dt;
bool g_host_built = false;
thread_local int g_last_hip = 0;

void build_host_once()
{
  if (!g_host_built) {
    build_tables(&g_hm, &g_dt);
    g_host_built = true;
  }
}

}  // namespace

int record_hip(hipError_t e)
{
  if (e == hipSuccess)
    return TK_OK;
  g_last_hip = (int)e;
  return TK_EDEVICE;
}

int after_launch()
{
  return record_hip(hipGetLastError());
}

int init_device(DevState **out)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    if (getenv("TK_DEBUG"))
      fprintf(stderr, "tkyber: hipGetDevice failed: %d\n", (int)e);
    record_hip(e);
    return TK_ENODEV;
  }
  if (dev < 0 || dev >= kMaxDevices)
    return TK_ENODEV;
  std::lock_guard<std::mutex> lk(g_mu);
  DevState &st = g_dev[dev];
  if (!st.tables) {
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) {
      if (getenv("TK_DEBUG"))
        fprintf(stderr, "tkyber: hipGetDeviceProperties failed: %d\n", (int)e);
      record_hip(e);
      return TK_ENODEV;
    }
    if (strstr(prop.gcnArchName, "gfx950") == nullptr) {
      if (getenv("TK_DEBUG"))
        fprintf(stderr, "tkyber: device %d is '%s', not gfx950\n", dev, prop.gcnArchName);
      return TK_ENODEV;
    }
    build_host_once();
    DeviceTables *d = nullptr;
    if (hipMalloc(&d, sizeof(DeviceTables)) != hipSuccess)
      return TK_ENOMEM;
    if ((e = hipMemcpy(d, &g_dt, sizeof(DeviceTables), hipMemcpyHostToDevice)) != hipSuccess) {
      (void)hipFree(d);
      return record_hip(e);
    }
    st.ncu = prop.multiProcessorCount;
    st.device = dev;
    int rc = noalloc_init(&st);
    if (rc) {
      (void)hipFree(d);
      return rc;
    }
    st.tables = d;
  }
  *out = &st;
  return TK_OK;
}

int device_ready(DevState **out)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return TK_ENODEV;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_dev[dev].tables)
    return TK_EINVAL;
  *out = &g_dev[dev];
  return TK_OK;
}

const HostMatrices &host_matrices()
{
  std::lock_guard<std::mutex> lk(g_mu);
  build_host_once();
  return g_hm;
}

const DeviceTables &host_device_tables()
{
  std::lock_guard<std::mutex> lk(g_mu);
  build_host_once();
  return g_dt;
}

bool aligned16(const void *p) { return p && ((uintptr_t)p & 15u) == 0; }

bool partial_overlap(const void *a, const void *b, size_t bytes)
{
  if (a == b)
    return false;
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + bytes && y < x + bytes;
}

int resolve_method(int method, int form, bool inverse, int *out_method, int *out_form)
{
  if (form == TK_FORM_DEFAULT)
    form = inverse ? TK_FORM_CENTERED : TK_FORM_CANONICAL;
  if (form < TK_FORM_CANONICAL || form > TK_FORM_REF)
    return TK_EINVAL;
  if (method == TK_METHOD_AUTO)
    method = form == TK_FORM_REF ? TK_METHOD_BUTTERFLY : TK_METHOD_MFMA;
  if (method == TK_METHOD_MFMA)
    method = TK_METHOD_MFMA16;
  if (method < TK_METHOD_BUTTERFLY || method > TK_METHOD_MFMA32)
    return TK_EINVAL;
  if (form == TK_FORM_REF && method != TK_METHOD_BUTTERFLY)
    return TK_ENOTSUP;
  *out_method = method;
  *out_form = form;
  return TK_OK;
}

int device_form(int form)
{
  switch (form) {
    case TK_FORM_CANONICAL: return kFormCanon;
    case TK_FORM_CENTERED: return kFormCentred;
    default: return kFormRef;
  }
}

// A fixed table rather than a map: launches after the first of each kernel
// allocate nothing (the no-allocation entry point relies on it).
int resident_per_cu(const void *kernel, int block_threads)
{
  struct Entry {
    const void *kernel;
    int block, n;
  };
  constexpr int kEntries = 1024;
  static std::mutex mu;
  static Entry cache[kEntries];
  static int used = 0;
  std::lock_guard<std::mutex> lk(mu);
  for (int i = 0; i < used; i++)
    if (cache[i].kernel == kernel && cache[i].block == block_threads)
      return cache[i].n;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block_threads, 0) != hipSuccess ||
      n < 1)
    n = 1;
  if (used < kEntries)
    cache[used++] = Entry{kernel, block_threads, n};
  return n;
}

int transform(bool inverse, int16_t *out, const int16_t *in, size_t npoly, int method, int form,
              void *stream)
{
  int m = 0, f = 0;
  int rc = resolve_method(method, form, inverse, &m, &f);
  if (rc)
    return rc;
  if (npoly == 0)
    return TK_OK;
  if (!aligned16(out) || !aligned16(in) || npoly > kMaxPolys)
    return TK_EINVAL;
  if (partial_overlap(out, in, npoly * kN * sizeof(int16_t)))
    return TK_EINVAL;
  DevState *st = nullptr;
  if ((rc = init_device(&st)))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  if (m == TK_METHOD_BUTTERFLY)
    rc = launch_butterfly(inverse, device_form(f), out, in, npoly, st->tables, st->ncu, s);
  else
    rc = launch_mfma(m == TK_METHOD_MFMA32 ? 32 : 16, inverse, device_form(f), out, in, npoly,
                     st->tables,
                     LaunchShape{st->ncu, 4, m == TK_METHOD_MFMA32 ? kLinearOversub32 : kLinearOversub},
                     s);
  if (rc)
    return TK_ENOTSUP;
  return after_launch();
}

namespace {

// One buffer of a fused call: address, size and, for an output, the one
// input it may alias exactly (the kernel reads that input's row in the same
// lane before it writes the output row at the same address).
struct Buf {
  const void *p;
  size_t bytes;
  const void *may_alias;
};

bool bufs_ok(const Buf *out, int nout, const Buf *in, int nin)
{
  auto overlap = [](const Buf &a, const Buf &b) {
    const uintptr_t x = (uintptr_t)a.p, y = (uintptr_t)b.p;
    return a.p && b.p && x < y + b.bytes && y < x + a.bytes;
  };
  for (int i = 0; i < nout; i++) {
    if (!aligned16(out[i].p))
      return false;
    for (int j = i + 1; j < nout; j++)
      if (overlap(out[i], out[j]))
        return false;
    for (int j = 0; j < nin; j++)
      if (overlap(out[i], in[j]) && !(out[i].may_alias == in[j].p && out[i].p == in[j].p))
        return false;
  }
  for (int j = 0; j < nin; j++)
    if (in[j].p && !aligned16(in[j].p))
      return false;
  return true;
}

}  // namespace

int validate_fused(int k, int mode, const FusedIo &io)
{
  if (k < 2 || k > 4 || io.inst)
    return TK_EINVAL;
  if (io.ninst > kMaxPolys / 16)
    return TK_EINVAL;
  const size_t pb = io.ninst * kN * sizeof(int16_t), vb = pb * k, mb = vb * k;
  const int16_t *x = static_cast<const int16_t *>(io.x);
  switch (mode) {
    case kFusedMatvec: {
      if (!io.out0 || !io.mat || !x)
        return TK_EINVAL;
      const Buf out[1] = {{io.out0, vb, x}};
      const Buf in[2] = {{io.mat, mb, nullptr}, {x, vb, nullptr}};
      return bufs_ok(out, 1, in, 2) ? TK_OK : TK_EINVAL;
    }
    case kFusedKeygen: {
      if (!io.out0 || !io.out1 || !io.mat || !x || !io.add0)
        return TK_EINVAL;
      const Buf out[2] = {{io.out0, vb, io.add0}, {io.out1, vb, x}};
      const Buf in[3] = {{io.mat, mb, nullptr}, {x, vb, nullptr}, {io.add0, vb, nullptr}};
      return bufs_ok(out, 2, in, 3) ? TK_OK : TK_EINVAL;
    }
    case kFusedEncrypt: {
      if (!io.out0 || !io.out1 || !io.mat || !io.mat2 || !x || !io.add0 || !io.add1)
        return TK_EINVAL;
      if (io.add2 && io.add2 == io.out1)  // v may alias e2 only
        return TK_EINVAL;
      const Buf out[2] = {{io.out0, vb, io.add0}, {io.out1, pb, io.add1}};
      const Buf in[6] = {{io.mat, mb, nullptr}, {io.mat2, vb, nullptr}, {x, vb, nullptr},
                         {io.add0, vb, nullptr}, {io.add1, pb, nullptr}, {io.add2, pb, nullptr}};
      return bufs_ok(out, 2, in, 6) ? TK_OK : TK_EINVAL;
    }
    case kFusedDecrypt: {
      if (!io.out0 || !io.mat || !x || !io.add0)
        return TK_EINVAL;
      const Buf out[1] = {{io.out0, pb, io.add0}};
      const Buf in[3] = {{io.mat, vb, nullptr}, {x, vb, nullptr}, {io.add0, pb, nullptr}};
      return bufs_ok(out, 1, in, 3) ? TK_OK : TK_EINVAL;
    }
    default:
      return TK_EINVAL;
  }
}

void fused_ranges(int k, int mode, const FusedIo &io, size_t req, std::vector<ReqRange> &outs,
                  std::vector<ReqRange> &ins)
{
  const size_t pb = io.ninst * kN * sizeof(int16_t), vb = pb * k, mb = vb * k;
  auto add = [&](std::vector<ReqRange> &v, const void *p, size_t bytes) {
    if (p && bytes)
      v.push_back(ReqRange{(uintptr_t)p, (uintptr_t)p + bytes, req});
  };
  switch (mode) {
    case kFusedMatvec:
      add(outs, io.out0, vb);
      add(ins, io.mat, mb);
      add(ins, io.x, vb);
      break;
    case kFusedKeygen:
      add(outs, io.out0, vb);
      add(outs, io.out1, vb);
      add(ins, io.mat, mb);
      add(ins, io.x, vb);
      add(ins, io.add0, vb);
      break;
    case kFusedEncrypt:
      add(outs, io.out0, vb);
      add(outs, io.out1, pb);
      add(ins, io.mat, mb);
      add(ins, io.mat2, vb);
      add(ins, io.x, vb);
      add(ins, io.add0, vb);
      add(ins, io.add1, pb);
      add(ins, io.add2, pb);
      break;
    default:  // kFusedDecrypt
      add(outs, io.out0, pb);
      add(ins, io.mat, vb);
      add(ins, io.x, vb);
      add(ins, io.add0, pb);
      break;
  }
}

bool cross_request_overlap(std::vector<ReqRange> &outs, const std::vector<ReqRange> &ins)
{
  std::sort(outs.begin(), outs.end(),
            [](const ReqRange &a, const ReqRange &b) { return a.lo < b.lo; });
  // Outputs must be pairwise disjoint (within a request too: its own
  // validation rejects overlapping outputs).
  for (size_t i = 1; i < outs.size(); i++)
    if (outs[i].lo < outs[i - 1].hi)
      return true;
  // Disjoint and sorted by start, so sorted by end as well: the outputs an
  // input range meets are a contiguous run from the first that ends after
  // its start.
  for (const ReqRange &in : ins) {
    auto it = std::upper_bound(outs.begin(), outs.end(), in.lo,
                               [](uintptr_t v, const ReqRange &o) { return v < o.hi; });
    for (; it != outs.end() && it->lo < in.hi; ++it)
      if (it->req != in.req)
        return true;
  }
  return false;
}

int run_fused_checked(int k, int mode, const FusedIo &io, void *stream)
{
  if (k < 2 || k > 4)
    return TK_EINVAL;
  if (io.ninst == 0)
    return TK_OK;
  int rc = validate_fused(k, mode, io);
  if (rc)
    return rc;
  DevState *st = nullptr;
  if ((rc = init_device(&st)))
    return rc;
  if (launch_fused(k, mode, 0, io, st->tables, st->ncu, (hipStream_t)stream))
    return TK_ENOTSUP;
  return after_launch();
}

}  // namespace tk

using namespace tk;

extern "C" {

int tk_init(void)
{
  DevState *st = nullptr;
  return init_device(&st);
}

const char *tk_strerror(int status)
{
  switch (status) {
    case TK_OK: return "ok";
    case TK_EINVAL: return "invalid argument";
    case TK_ENOMEM: return "device allocation failed";
    case TK_EDEVICE: return "HIP runtime error";
    case TK_ENOTSUP: return "unsupported method/form combination";
    case TK_ENODEV: return "no gfx950 device";
    default: return "unknown status";
  }
}

int tk_last_hip_error(void) { return g_last_hip; }

const char *tk_version(void) { return "tensorkyber_amd 0.1 (gfx950)"; }

int ntt_batch(int16_t *polys, size_t npoly, void *stream)
{
  return transform(false, polys, polys, npoly, TK_METHOD_AUTO, TK_FORM_DEFAULT, stream);
}

int invntt_batch(int16_t *polys, size_t npoly, void *stream)
{
  return transform(true, polys, polys, npoly, TK_METHOD_AUTO, TK_FORM_DEFAULT, stream);
}

int ntt_batch_ex(int16_t *out, const int16_t *in, size_t npoly, int method, int form,
                 void *stream)
{
  return transform(false, out, in, npoly, method, form, stream);
}

int invntt_batch_ex(int16_t *out, const int16_t *in, size_t npoly, int method, int form,
                    void *stream)
{
  return transform(true, out, in, npoly, method, form, stream);
}

int basemul_batch_ex(int16_t *r, const int16_t *a, const int16_t *b, size_t npoly, int form,
                     void *stream)
{
  if (form == TK_FORM_DEFAULT)
    form = TK_FORM_REF;
  if (form < TK_FORM_CANONICAL || form > TK_FORM_REF)
    return TK_EINVAL;
  if (npoly == 0)
    return TK_OK;
  if (!aligned16(r) || !aligned16(a) || !aligned16(b) || npoly > kMaxPolys)
    return TK_EINVAL;
  const size_t bytes = npoly * kN * sizeof(int16_t);
  if (partial_overlap(r, a, bytes) || partial_overlap(r, b, bytes))
    return TK_EINVAL;
  DevState *st = nullptr;
  int rc = init_device(&st);
  if (rc)
    return rc;
  if (launch_basemul(device_form(form), r, a, b, npoly, st->tables, st->ncu,
                     (hipStream_t)stream))
    return TK_ENOTSUP;
  return after_launch();
}

int basemul_batch(int16_t *r, const int16_t *a, const int16_t *b, size_t npoly, void *stream)
{
  return basemul_batch_ex(r, a, b, npoly, TK_FORM_DEFAULT, stream);
}

int matvec_ntt_fused(int16_t *t, const int16_t *A_hat, const int16_t *s, int k, size_t ninst,
                     void *stream)
{
  // t may alias s exactly: each tile reads all of its s before writing t.
  FusedIo io{};
  io.out0 = t;
  io.mat = A_hat;
  io.x = s;
  io.ninst = ninst;
  return run_fused_checked(k, kFusedMatvec, io, stream);
}

int kpke_keygen_core(int16_t *t_hat, int16_t *s_hat, const int16_t *A_hat, const int16_t *s,
                     const int16_t *e, int k, size_t ninst, void *stream)
{
  FusedIo io{};
  io.out0 = t_hat;
  io.out1 = s_hat;
  io.mat = A_hat;
  io.x = s;
  io.add0 = e;
  io.ninst = ninst;
  return run_fused_checked(k, kFusedKeygen, io, stream);
}

int kpke_encrypt_core(int16_t *u, int16_t *v, const int16_t *At_hat, const int16_t *t_hat,
                      const int16_t *r, const int16_t *e1, const int16_t *e2, const int16_t *m,
                      int k, size_t ninst, void *stream)
{
  FusedIo io{};
  io.out0 = u;
  io.out1 = v;
  io.mat = At_hat;
  io.mat2 = t_hat;
  io.x = r;
  io.add0 = e1;
  io.add1 = e2;
  io.add2 = m;
  io.ninst = ninst;
  return run_fused_checked(k, kFusedEncrypt, io, stream);
}

int kpke_decrypt_core(int16_t *w, const int16_t *s_hat, const int16_t *u, const int16_t *v, int k,
                      size_t ninst, void *stream)
{
  FusedIo io{};
  io.out0 = w;
  io.mat = s_hat;
  io.x = u;
  io.add0 = v;
  io.ninst = ninst;
  return run_fused_checked(k, kFusedDecrypt, io, stream);
}

size_t tk_packed_poly_bytes(int bits)
{
  return (bits >= 2 && bits <= 4) ? (size_t)(kN / (8 / bits)) : 0;
}

int pack_small_batch(uint8_t *packed, const int16_t *in, size_t npoly, int bits, void *stream)
{
  const size_t ppb = tk_packed_poly_bytes(bits);
  if (ppb == 0)
    return TK_EINVAL;
  if (npoly == 0)
    return TK_OK;
  if (!aligned16(packed) || !aligned16(in) || npoly > kMaxPolys)
    return TK_EINVAL;
  const uintptr_t pp = (uintptr_t)packed, ip = (uintptr_t)in;
  if (pp < ip + npoly * kN * sizeof(int16_t) && ip < pp + npoly * ppb)
    return TK_EINVAL;
  DevState *st = nullptr;
  int rc = init_device(&st);
  if (rc)
    return rc;
  if (launch_pack_small(bits, packed, in, npoly, st->ncu, (hipStream_t)stream))
    return TK_ENOTSUP;
  return after_launch();
}

int ntt_batch_packed(int16_t *out, const uint8_t *packed, size_t npoly, int bits, int method,
                     int form, void *stream)
{
  const size_t ppb = tk_packed_poly_bytes(bits);
  if (ppb == 0)
    return TK_EINVAL;
  int m = 0, f = 0;
  int rc = resolve_method(method == TK_METHOD_AUTO ? TK_METHOD_MFMA : method, form, false, &m, &f);
  if (rc)
    return rc;
  if (m == TK_METHOD_BUTTERFLY)
    return TK_ENOTSUP;
  if (npoly == 0)
    return TK_OK;
  if (!aligned16(out) || !aligned16(packed) || npoly > kMaxPolys)
    return TK_EINVAL;
  const uintptr_t op = (uintptr_t)out, pp = (uintptr_t)packed;
  if (op < pp + npoly * ppb && pp < op + npoly * kN * sizeof(int16_t))
    return TK_EINVAL;
  DevState *st = nullptr;
  if ((rc = init_device(&st)))
    return rc;
  if (launch_mfma_packed(m == TK_METHOD_MFMA32 ? 32 : 16, bits, device_form(f), out, packed, npoly,
                         st->tables, LaunchShape{st->ncu, packed_waves(bits), kPackedOversub}, (hipStream_t)stream))
    return TK_ENOTSUP;
  return after_launch();
}

int matvec_ntt_fused_packed(int16_t *t, const int16_t *A_hat, const uint8_t *s_packed, int k,
                            size_t ninst, int bits, void *stream)
{
  const size_t ppb = tk_packed_poly_bytes(bits);
  if (ppb == 0 || k < 2 || k > 4)
    return TK_EINVAL;
  if (ninst == 0)
    return TK_OK;
  if (!aligned16(t) || !aligned16(A_hat) || !aligned16(s_packed) || ninst > kMaxPolys / 16)
    return TK_EINVAL;
  const size_t vb = ninst * k * kN * sizeof(int16_t), mb = vb * k, sb = ninst * k * ppb;
  const uintptr_t tp = (uintptr_t)t, ap = (uintptr_t)A_hat, sp = (uintptr_t)s_packed;
  if ((tp < ap + mb && ap < tp + vb) || (tp < sp + sb && sp < tp + vb))
    return TK_EINVAL;
  DevState *st = nullptr;
  int rc = init_device(&st);
  if (rc)
    return rc;
  FusedIo io{};
  io.out0 = t;
  io.mat = A_hat;
  io.x = s_packed;
  io.ninst = ninst;
  if (launch_fused(k, kFusedMatvec, bits, io, st->tables, st->ncu, (hipStream_t)stream))
    return TK_ENOTSUP;
  return after_launch();
}

int tk_host_tables(int16_t *zetas, int16_t *fwd1, int16_t *fwd2, int16_t *inv2, int16_t *inv1)
{
  const HostMatrices &hm = host_matrices();
  if (zetas)
    memcpy(zetas, hm.zetas, sizeof(hm.zetas));
  if (fwd1)
    memcpy(fwd1, hm.fwd1, sizeof(hm.fwd1));
  if (fwd2)
    memcpy(fwd2, hm.fwd2, sizeof(hm.fwd2));
  if (inv2)
    memcpy(inv2, hm.inv2, sizeof(hm.inv2));
  if (inv1)
    memcpy(inv1, hm.inv1, sizeof(hm.inv1));
  return TK_OK;
}

int tk_host_mfma_operands(int shape, int stage, uint32_t *ops, int32_t *bias)
{
  if ((shape != 16 && shape != 32) || stage < kFwd1 || stage > kInv1 + 3)
    return TK_EINVAL;
  if (stage > kInv1 && shape != 16)
    return TK_EINVAL;
  const DeviceTables &dt = host_device_tables();
  if (stage > kInv1) {  // packed first pass, 2 + (stage - 4) bits
    if (ops)
      memcpy(ops, dt.m16_fwd1pk[stage - kInv1 - 1], sizeof(dt.m16_fwd1pk[0]));
    if (bias)
      memcpy(bias, dt.bias_fwd1pk, sizeof(dt.bias_fwd1pk));
    return TK_OK;
  }
  const void *src = nullptr;
  size_t n = 0;
  const int32_t *b = nullptr;
  size_t nb = 0;
  switch (stage) {
    case kFwd1:
      src = shape == 16 ? (const void *)dt.m16_fwd1 : (const void *)dt.m32_fwd1;
      n = shape == 16 ? sizeof(dt.m16_fwd1) : sizeof(dt.m32_fwd1);
      b = &dt.bias_fwd1[0][0];
      nb = sizeof(dt.bias_fwd1);
      break;
    case kFwd2:
      src = shape == 16 ? (const void *)dt.m16_fwd2 : (const void *)dt.m32_fwd2;
      n = shape == 16 ? sizeof(dt.m16_fwd2) : sizeof(dt.m32_fwd2);
      b = &dt.bias_fwd2[0][0];
      nb = sizeof(dt.bias_fwd2);
      break;
    case kInv2:
      src = shape == 16 ? (const void *)dt.m16_inv2 : (const void *)dt.m32_inv2;
      n = shape == 16 ? sizeof(dt.m16_inv2) : sizeof(dt.m32_inv2);
      b = &dt.bias_inv2[0][0];
      nb = sizeof(dt.bias_inv2);
      break;
    default:
      src = shape == 16 ? (const void *)dt.m16_inv1 : (const void *)dt.m32_inv1;
      n = shape == 16 ? sizeof(dt.m16_inv1) : sizeof(dt.m32_inv1);
      b = &dt.bias_inv1[0][0];
      nb = sizeof(dt.bias_inv1);
      break;
  }
  if (ops)
    memcpy(ops, src, n);
  if (bias)
    memcpy(bias, b, nb);
  return TK_OK;
}

}  // extern "C"
