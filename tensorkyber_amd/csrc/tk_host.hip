// Host-resident batches: the reference's own calling shape is ntt(r) on host
// memory (ref/ntt.h:14-15, called per poly from ntt_bench.cu:947-956).
//
// Pinned host buffers (hipHostMalloc, torch pin_memory) are transformed by
// the kernel itself, reading its input and writing its output over PCIe:
// both link directions carry data at once and nothing is staged.  Copy
// engines serialise opposite-direction copies on this device (two
// 512 MiB copies on separate streams move 29 GB/s each way, one alone 57).
// A buffer qualifies only if the whole byte range lies inside one pinned
// allocation (hipMemGetAddressRange), so the kernel never touches pageable
// memory.
//
// Other buffers stream through the device in chunks: chunk c is copied
// host -> device, transformed in place and copied back on stream c % depth,
// so with depth >= 2 the copies of one chunk overlap the kernel and the
// copies of the others.  Pageable buffers go through the HIP runtime's own
// staging and are slower.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tk_state.h"
#include "tkyber.h"

struct tk_host_pipeline {
  int device = -1;
  size_t chunk = 0;
  int depth = 0;
  int16_t *dev[TK_HOST_MAX_DEPTH] = {};
  hipStream_t stream[TK_HOST_MAX_DEPTH] = {};
  std::mutex mu;  // one run at a time per pipeline
};

namespace {

void pipeline_free(tk_host_pipeline *p)
{
  for (int i = 0; i < TK_HOST_MAX_DEPTH; i++) {
    if (p->stream[i])
      (void)hipStreamDestroy(p->stream[i]);
    if (p->dev[i])
      (void)hipFree(p->dev[i]);
  }
  delete p;
}

// The device address of [p, p + bytes) if that range lies inside one pinned
// host allocation, else nullptr.  The queries fail for pageable memory and
// leave the runtime's last error set; it is cleared, so that the next launch
// check does not report it.
void *pinned_span(const void *p, size_t bytes)
{
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (a.type != hipMemoryTypeHost || !a.devicePointer)
    return nullptr;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)a.devicePointer) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  const uintptr_t b = (uintptr_t)base, d = (uintptr_t)a.devicePointer;
  if (d < b || d - b > size || size - (d - b) < bytes)
    return nullptr;
  return a.devicePointer;
}

// Default pipeline per device for the ntt_batch_host convenience calls.
std::mutex g_default_mu;
tk_host_pipeline *g_default[tk::kMaxDevices] = {};

int default_pipeline(tk_host_pipeline **out)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= tk::kMaxDevices)
    return TK_ENODEV;
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default[dev]) {
    int rc = tk_host_pipeline_create(&g_default[dev], TK_HOST_DEFAULT_CHUNK, 3);
    if (rc)
      return rc;
  }
  *out = g_default[dev];
  return TK_OK;
}

}  // namespace

extern "C" {

int tk_host_pipeline_create(tk_host_pipeline **out, size_t chunk_polys, int depth)
{
  if (!out || chunk_polys == 0 || chunk_polys > tk::kMaxPolys || depth < 1 ||
      depth > TK_HOST_MAX_DEPTH)
    return TK_EINVAL;
  *out = nullptr;
  tk::DevState *st = nullptr;
  int rc = tk::init_device(&st);
  if (rc)
    return rc;
  tk_host_pipeline *p = new tk_host_pipeline();
  p->device = st->device;
  p->chunk = chunk_polys;
  p->depth = depth;
  for (int i = 0; i < depth; i++) {
    if (hipMalloc(&p->dev[i], chunk_polys * tk::kN * sizeof(int16_t)) != hipSuccess) {
      pipeline_free(p);
      return TK_ENOMEM;
    }
    if ((rc = tk::record_hip(hipStreamCreateWithFlags(&p->stream[i], hipStreamNonBlocking)))) {
      pipeline_free(p);
      return rc;
    }
  }
  *out = p;
  return TK_OK;
}

int tk_host_pipeline_run(tk_host_pipeline *p, int16_t *out, const int16_t *in, size_t npoly,
                         int direction, int method, int form)
{
  if (!p || (direction != 0 && direction != 1))
    return TK_EINVAL;
  if (npoly == 0)
    return TK_OK;
  if (!out || !in || npoly > tk::kMaxPolys)
    return TK_EINVAL;
  const size_t bytes = npoly * tk::kN * sizeof(int16_t);
  if (tk::partial_overlap(out, in, bytes))
    return TK_EINVAL;
  // Validate method and form before anything is queued.
  int m = 0, f = 0;
  int rc = tk::resolve_method(method, form, direction != 0, &m, &f);
  if (rc)
    return rc;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != p->device)
    return TK_ENODEV;
  std::lock_guard<std::mutex> lk(p->mu);
  void *din = pinned_span(in, bytes), *dout = din ? pinned_span(out, bytes) : nullptr;
  if (din && dout) {
    rc = tk::transform(direction != 0, static_cast<int16_t *>(dout),
                       static_cast<const int16_t *>(din), npoly, m, f, p->stream[0]);
    const int src = tk::record_hip(hipStreamSynchronize(p->stream[0]));
    return rc ? rc : src;
  }
  const size_t nchunks = (npoly + p->chunk - 1) / p->chunk;
  for (size_t c = 0; c < nchunks && rc == TK_OK; c++) {
    const int slot = (int)(c % (size_t)p->depth);
    const size_t first = c * p->chunk;
    const size_t n = npoly - first < p->chunk ? npoly - first : p->chunk;
    const size_t off = first * tk::kN, nb = n * tk::kN * sizeof(int16_t);
    hipStream_t s = p->stream[slot];
    // Same-stream order keeps this chunk's copy-in behind the previous
    // chunk's copy-out of the same device buffer.
    if ((rc = tk::record_hip(hipMemcpyAsync(p->dev[slot], in + off, nb, hipMemcpyHostToDevice, s))))
      break;
    if ((rc = tk::transform(direction != 0, p->dev[slot], p->dev[slot], n, m, f, s)))
      break;
    rc = tk::record_hip(hipMemcpyAsync(out + off, p->dev[slot], nb, hipMemcpyDeviceToHost, s));
  }
  for (int i = 0; i < p->depth; i++) {
    const int src = tk::record_hip(hipStreamSynchronize(p->stream[i]));
    if (rc == TK_OK)
      rc = src;
  }
  return rc;
}

int tk_host_pipeline_destroy(tk_host_pipeline *p)
{
  if (!p)
    return TK_EINVAL;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    for (int i = 0; i < p->depth; i++)
      (void)hipStreamSynchronize(p->stream[i]);
  }
  pipeline_free(p);
  return TK_OK;
}

int ntt_batch_host(int16_t *polys, size_t npoly)
{
  if (npoly == 0)
    return TK_OK;
  if (!polys)
    return TK_EINVAL;
  tk_host_pipeline *p = nullptr;
  int rc = default_pipeline(&p);
  return rc ? rc : tk_host_pipeline_run(p, polys, polys, npoly, 0, TK_METHOD_AUTO, TK_FORM_DEFAULT);
}

int invntt_batch_host(int16_t *polys, size_t npoly)
{
  if (npoly == 0)
    return TK_OK;
  if (!polys)
    return TK_EINVAL;
  tk_host_pipeline *p = nullptr;
  int rc = default_pipeline(&p);
  return rc ? rc : tk_host_pipeline_run(p, polys, polys, npoly, 1, TK_METHOD_AUTO, TK_FORM_DEFAULT);
}

}  // extern "C"
