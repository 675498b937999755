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
// Other buffers (pageable memory, which the device cannot address) are
// staged by the pipeline itself: worker threads copy chunk c into pinned
// staging buffer c % depth, the kernel transforms it in place over PCIe on
// stream c % depth, and the result is copied back once that stream is done,
// while later chunks are being staged and transformed.  The HIP runtime's
// own pageable copies are serial and reach a quarter of that rate.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "tk_state.h"
#include "tkyber.h"

namespace {

// A fixed set of threads running one job at a time, each thread on its own
// share (the caller runs share 0): the copies between a caller's pageable
// buffers and a pipeline's pinned staging.
class CopyPool {
 public:
  // Throws if a thread cannot be started, after joining those that were.
  explicit CopyPool(int n) : n_(n)
  {
    try {
      for (int i = 1; i < n_; i++)
        th_.emplace_back([this, i] { loop(i); });
    } catch (...) {
      stop_all();
      throw;
    }
  }
  ~CopyPool() { stop_all(); }
  int size() const { return n_; }
  // f(0) .. f(size() - 1); returns when every share is done.
  void run(const std::function<void(int)> &f)
  {
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &f;
      left_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    f(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return left_ == 0; });
    job_ = nullptr;
  }

 private:
  void stop_all()
  {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_)
      t.join();
    th_.clear();
  }
  void loop(int i)
  {
    unsigned seen = 0;
    for (;;) {
      const std::function<void(int)> *f = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_)
          return;
        seen = gen_;
        f = job_;
      }
      (*f)(i);
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0)
        done_.notify_one();
    }
  }
  int n_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *job_ = nullptr;
  int left_ = 0;
  unsigned gen_ = 0;
  bool stop_ = false;
};

// memcpy of `bytes` split into page-aligned shares over the pool.
void parallel_copy(CopyPool &pool, void *dst, const void *src, size_t bytes)
{
  const size_t n = (size_t)pool.size();
  const size_t share = ((bytes + n - 1) / n + 4095) & ~(size_t)4095;
  pool.run([&](int i) {
    const size_t lo = std::min(bytes, (size_t)i * share), hi = std::min(bytes, lo + share);
    if (hi > lo)
      memcpy(static_cast<char *>(dst) + lo, static_cast<const char *>(src) + lo, hi - lo);
  });
}

constexpr int kCopyThreads = 8;

}  // namespace

struct tk_host_pipeline {
  int device = -1;
  size_t chunk = 0;
  int depth = 0;
  int16_t *stage[TK_HOST_MAX_DEPTH] = {};   // pinned host staging, chunk polys each
  int16_t *dstage[TK_HOST_MAX_DEPTH] = {};  // their device addresses
  hipStream_t stream[TK_HOST_MAX_DEPTH] = {};
  CopyPool *pool = nullptr;
  std::mutex mu;  // one run at a time per pipeline
};

namespace {

void pipeline_free(tk_host_pipeline *p)
{
  for (int i = 0; i < TK_HOST_MAX_DEPTH; i++) {
    if (p->stream[i])
      (void)hipStreamDestroy(p->stream[i]);
    if (p->stage[i])
      (void)hipHostFree(p->stage[i]);
  }
  delete p->pool;
  delete p;
}

// Where the bytes [p, p + bytes) live:
//   kPinned  inside one pinned host allocation; *dev is their device address;
//   kDevice  device memory, which the host copies of the staged path cannot
//            touch (the run is rejected);
//   kOther   anything else, taken to be host memory the device cannot
//            address (pageable), or a pinned range that overruns its
//            allocation.
// The queries fail for pageable memory and leave the runtime's last error
// set; it is cleared, so that the next launch check does not report it.
enum SpanKind { kOther, kPinned, kDevice };

SpanKind host_span(const void *p, size_t bytes, void **dev)
{
  *dev = nullptr;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return kOther;
  }
  if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeArray)
    return kDevice;
  if (a.type != hipMemoryTypeHost || !a.devicePointer)
    return kOther;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)a.devicePointer) != hipSuccess) {
    (void)hipGetLastError();
    return kOther;
  }
  const uintptr_t b = (uintptr_t)base, d = (uintptr_t)a.devicePointer;
  if (d < b || d - b > size || size - (d - b) < bytes)
    return kOther;
  *dev = a.devicePointer;
  return kPinned;
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
    int rc = tk_host_pipeline_create(&g_default[dev], TK_HOST_DEFAULT_CHUNK,
                                     TK_HOST_DEFAULT_DEPTH);
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
  tk_host_pipeline *p = new (std::nothrow) tk_host_pipeline();
  if (!p)
    return TK_ENOMEM;
  p->device = st->device;
  p->chunk = chunk_polys;
  p->depth = depth;
  for (int i = 0; i < depth; i++) {
    if (hipHostMalloc(reinterpret_cast<void **>(&p->stage[i]),
                      chunk_polys * tk::kN * sizeof(int16_t), hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&p->dstage[i]), p->stage[i], 0) !=
            hipSuccess) {
      (void)hipGetLastError();
      pipeline_free(p);
      return TK_ENOMEM;
    }
    if ((rc = tk::record_hip(hipStreamCreateWithFlags(&p->stream[i], hipStreamNonBlocking)))) {
      pipeline_free(p);
      return rc;
    }
  }
  try {
    p->pool = new CopyPool(std::max(1u, std::min((unsigned)kCopyThreads,
                                                 std::thread::hardware_concurrency())));
  } catch (...) {
    pipeline_free(p);
    return TK_ENOMEM;
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
  void *din = nullptr, *dout = nullptr;
  const SpanKind kin = host_span(in, bytes, &din), kout = host_span(out, bytes, &dout);
  if (kin == kDevice || kout == kDevice)
    return TK_EINVAL;
  std::lock_guard<std::mutex> lk(p->mu);
  if (kin == kPinned && kout == kPinned) {
    rc = tk::transform(direction != 0, static_cast<int16_t *>(dout),
                       static_cast<const int16_t *>(din), npoly, m, f, p->stream[0]);
    const int src = tk::record_hip(hipStreamSynchronize(p->stream[0]));
    return rc ? rc : src;
  }
  // Staged: chunk c in staging buffer c % depth on stream c % depth; the
  // result of chunk c - depth is copied out before that buffer is refilled.
  const size_t nchunks = (npoly + p->chunk - 1) / p->chunk;
  auto span = [&](size_t c, size_t *off, size_t *nb) {
    const size_t first = c * p->chunk, n = std::min(p->chunk, npoly - first);
    *off = first * tk::kN;
    *nb = n * tk::kN * sizeof(int16_t);
    return n;
  };
  auto retire = [&](size_t c) {
    const int slot = (int)(c % (size_t)p->depth);
    size_t off, nb;
    span(c, &off, &nb);
    const int src = tk::record_hip(hipStreamSynchronize(p->stream[slot]));
    if (src == TK_OK)
      parallel_copy(*p->pool, out + off, p->stage[slot], nb);
    return src;
  };
  size_t launched = 0;
  for (size_t c = 0; c < nchunks && rc == TK_OK; c++) {
    const int slot = (int)(c % (size_t)p->depth);
    if (c >= (size_t)p->depth && (rc = retire(c - p->depth)))
      break;
    size_t off, nb;
    const size_t n = span(c, &off, &nb);
    parallel_copy(*p->pool, p->stage[slot], in + off, nb);
    if ((rc = tk::transform(direction != 0, p->dstage[slot], p->dstage[slot], n, m, f,
                            p->stream[slot])))
      break;
    launched = c + 1;
  }
  const size_t first_open = launched > (size_t)p->depth ? launched - p->depth : 0;
  for (size_t c = first_open; c < launched; c++) {
    const int rrc = rc == TK_OK ? retire(c) : tk::record_hip(hipStreamSynchronize(p->stream[c % p->depth]));
    if (rc == TK_OK)
      rc = rrc;
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
