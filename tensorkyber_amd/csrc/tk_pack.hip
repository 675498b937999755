// Packing of small-coefficient polys into the BITS-bit form that
// ntt_batch_packed and matvec_ntt_fused_packed read: BITS-bit two's-complement
// fields, 8/BITS per byte, coefficient i in byte i/(8/BITS) at bit
// BITS*(i%(8/BITS)) -- the reference's pack_small (ntt_bench.cu:749-757),
// which keeps the low BITS bits of each value.
//
// One lane writes one packed dword: 16 coefficients (32 input bytes) for
// BITS = 2, 8 (16 bytes) for BITS = 3 or 4.  Dwords tile both buffers
// contiguously (a poly is a whole number of them), so consecutive lanes read
// and write consecutive addresses.
#include "tk_grid.h"
#include "tk_launch.h"
#include "tk_mfma_core.h"

namespace tk {
namespace {

template <int BITS>
__global__ __launch_bounds__(256) void k_pack_small(uint8_t *__restrict__ out,
                                                    const int16_t *__restrict__ in, size_t ndw)
{
  constexpr int PER = 8 / BITS;  // fields per byte
  constexpr int CPD = 4 * PER;   // coefficients per output dword
  constexpr uint32_t kMask = (1u << BITS) - 1;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < ndw;
       q += (size_t)gridDim.x * blockDim.x) {
    const v4i *src = reinterpret_cast<const v4i *>(in + q * CPD);
    uint32_t w = 0;
#pragma unroll
    for (int h = 0; h < CPD / 8; h++) {
      const v4i x = ld_stream(src + h);
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int i = 8 * h + e;
        const uint32_t v = ((uint32_t)x[e >> 1] >> (16 * (e & 1))) & kMask;
        w |= v << (8 * (i / PER) + BITS * (i % PER));
      }
    }
    __builtin_nontemporal_store(w, reinterpret_cast<uint32_t *>(out) + q);
  }
}

}  // namespace

int launch_pack_small(int bits, uint8_t *out, const int16_t *in, size_t npoly, int ncu,
                      hipStream_t stream)
{
  // The grid covers the batch (one dword per lane); the loop only takes
  // over beyond the largest grid (2^32 - 1 lanes: 2^28 polys at 2 bits).  On
  // MI355X this packs 2^20 polys into 2 bits in 99 us, against 116 us for a
  // grid of 16 workgroups per CU looping over the batch (3 bits: 113-114 us
  // either way).
  (void)ncu;
  auto go = [&](auto kernel, int ppb) {
    const size_t ndw = npoly * (size_t)(ppb / 4);
    hipLaunchKernelGGL(kernel, dim3(stream_grid(ndw, 256)), dim3(256), 0, stream, out, in, ndw);
    return 0;
  };
  switch (bits) {
    case 2: return go(k_pack_small<2>, packed_poly_bytes<2>());
    case 3: return go(k_pack_small<3>, packed_poly_bytes<3>());
    case 4: return go(k_pack_small<4>, packed_poly_bytes<4>());
    default: return -1;
  }
}

}  // namespace tk
