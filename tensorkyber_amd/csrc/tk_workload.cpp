// Deterministic synthetic inputs (host only): counter-based SplitMix64, so any
// element can be generated independently and the work splits across threads.
#include "tkyber_workload.h"

#include <algorithm>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t key_of(uint64_t seed, uint64_t stream) { return mix64(seed) ^ stream; }

inline uint64_t draw_k(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * kGolden); }

inline uint32_t below_k(uint64_t key, uint64_t i, uint32_t n)
{
  return (uint32_t)(((draw_k(key, i) >> 32) * (uint64_t)n) >> 32);
}

template <class F>
void parallel_for(size_t n, F &&f)
{
  unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < (1u << 16) || hw == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  size_t chunk = (n + hw - 1) / hw;
  for (unsigned t = 0; t < hw; t++) {
    size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e)
      break;
    th.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto &x : th)
    x.join();
}

constexpr int kQ = 3329;

}  // namespace

extern "C" {

uint64_t tk_wl_draw(uint64_t seed, uint64_t stream, uint64_t index)
{
  return draw_k(key_of(seed, stream), index);
}

uint32_t tk_wl_below(uint64_t seed, uint64_t stream, uint64_t index, uint32_t n)
{
  return below_k(key_of(seed, stream), index, n);
}

void tk_wl_uniform(int16_t *out, size_t n, uint64_t seed, uint64_t stream, int lo, int hi)
{
  const uint64_t key = key_of(seed, stream);
  const uint32_t span = (uint32_t)(hi - lo);
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; i++)
      out[i] = (int16_t)(lo + (int)below_k(key, i, span));
  });
}

void tk_wl_edge_mix(int16_t *out, size_t npoly, uint64_t seed)
{
  const uint64_t kc = key_of(seed, TK_WL_COEF), kcat = key_of(seed, TK_WL_CAT),
                 kpos = key_of(seed, TK_WL_POS);
  parallel_for(npoly, [&](size_t b, size_t e) {
    for (size_t p = b; p < e; p++) {
      int16_t *r = out + p * 256;
      const uint32_t cat = below_k(kcat, p, 100);
      if (cat < 70) {
        for (int j = 0; j < 256; j++)
          r[j] = (int16_t)((int)below_k(kc, p * 256 + j, 2 * (kQ - 1) + 1) - (kQ - 1));
      } else {
        const int16_t v = cat < 80 ? 0 : cat < 90 ? (int16_t)(kQ - 1) : cat < 95 ? (int16_t)(1 - kQ) : 0;
        for (int j = 0; j < 256; j++)
          r[j] = v;
        if (cat >= 95)
          r[below_k(kpos, p, 256)] = 1;
      }
    }
  });
}

void tk_wl_cbd2(int16_t *out, size_t n, uint64_t seed, uint64_t stream)
{
  const uint64_t key = key_of(seed, stream);
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; i++) {
      const uint64_t v = draw_k(key, i) >> 60;
      out[i] = (int16_t)((int)((v & 1) + ((v >> 1) & 1)) - (int)(((v >> 2) & 1) + ((v >> 3) & 1)));
    }
  });
}

size_t tk_wl_ragged_layout(tk_wl_request *reqs, size_t max_req, size_t npoly_total, uint64_t seed,
                           size_t *pool_polys)
{
  const uint64_t kr = key_of(seed, TK_WL_REQ), kg = key_of(seed, TK_WL_GAP);
  size_t done = 0, pool = 0, nreq = 0;
  while (done < npoly_total) {
    const uint16_t kind = below_k(kr, 3 * nreq + 0, 4) == 0 ? 1 : 0;
    const uint16_t k = (uint16_t)(2 + below_k(kr, 3 * nreq + 1, 3));
    const uint32_t obj = kind ? (uint32_t)k * k : k;
    const uint32_t count = 1 + below_k(kr, 3 * nreq + 2, 4096 / obj);
    size_t np = (size_t)count * obj;
    if (np > npoly_total - done)
      np = npoly_total - done;
    pool += below_k(kg, nreq, 64);
    if (reqs && nreq < max_req) {
      reqs[nreq].offset = pool;
      reqs[nreq].npoly = (uint32_t)np;
      reqs[nreq].k = k;
      reqs[nreq].kind = kind;
    }
    pool += np;
    done += np;
    nreq++;
  }
  if (pool_polys)
    *pool_polys = pool;
  return nreq;
}

uint64_t tk_wl_canon_digest(const int16_t *polys, size_t ncoef)
{
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < ncoef; i++) {
    int c = polys[i] % kQ;
    if (c < 0)
      c += kQ;
    h ^= (uint64_t)(c & 0xff);
    h *= 0x100000001b3ull;
    h ^= (uint64_t)(c >> 8);
    h *= 0x100000001b3ull;
  }
  return h;
}

}  // extern "C"
