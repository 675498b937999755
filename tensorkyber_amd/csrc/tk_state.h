// Per-device library state and the validation helpers shared by the C ABI
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "tk_internal.h"
#include "tk_launch.h"

namespace tk {

constexpr int kMaxDevices = 64;
// Upper bound on one batch: 2^36 polys (32 TiB) keeps every index in size_t
// arithmetic far from overflow.
constexpr size_t kMaxPolys = (size_t)1 << 36;

struct DevState {
  int device = -1;
  int ncu = 0;
  DeviceTables *tables = nullptr;
  // Device address of this device's copy of the no-allocation entry
  // point's static scratch array (tk_noalloc.hip).
  void *noalloc_scratch = nullptr;
};

// Uploads the tables to the current device on first use.  Thread-safe.
int init_device(DevState **out);
// The current device's state if init_device has run on it (TK_EINVAL if
// not), without initialising anything: no allocation.
int device_ready(DevState **out);
// Per-device set-up of the no-allocation entry point, from init_device with
// the device current: the scratch array's device address and the launch
// shapes it will use.
int noalloc_init(DevState *st);
int record_hip(hipError_t e);
int after_launch();
bool aligned16(const void *p);
bool partial_overlap(const void *a, const void *b, size_t bytes);
int resolve_method(int method, int form, bool inverse, int *out_method, int *out_form);
int device_form(int form);
// Validated batch transform on device buffers (the body of ntt_batch_ex /
// invntt_batch_ex).
int transform(bool inverse, int16_t *out, const int16_t *in, size_t npoly, int method, int form,
              void *stream);
// Argument rules of a fused call (k, required buffers, alignment, the one
// exact alias each output may have, no other overlap); TK_OK or TK_EINVAL.
int validate_fused(int k, int mode, const FusedIo &io);
// Byte range [lo, hi) of a buffer of request `req` of a plan.
struct ReqRange {
  uintptr_t lo, hi;
  size_t req;
};
// True if an output range overlaps any other output range, or an input range
// of another request.  Ranges of one request are checked by that request's
// own validation (an output may alias its own input exactly).  Sorts `outs`.
bool cross_request_overlap(std::vector<ReqRange> &outs, const std::vector<ReqRange> &ins);
// The output and input byte ranges of a fused call's buffers (non-null ones),
// tagged with `req`.
void fused_ranges(int k, int mode, const FusedIo &io, size_t req, std::vector<ReqRange> &outs,
                  std::vector<ReqRange> &ins);
// validate_fused, then one launch on stream.
int run_fused_checked(int k, int mode, const FusedIo &io, void *stream);
const HostMatrices &host_matrices();
const DeviceTables &host_device_tables();

}  // namespace tk
