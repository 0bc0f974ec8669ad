// gen.hip -- counter-based synthetic-input generators (include/lnt_testgen.h).
// Integer-only and exact-IEEE arithmetic, matching the CPU oracle bit for bit.
#include "../../include/lnt_testgen.h"
#include <string>

#include "lnt_internal.h"

namespace lnt {
namespace {

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline uint64_t rand_u64(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

__device__ inline uint64_t feistel(uint64_t x, int bits, uint64_t seed) {
  int half = bits / 2;
  uint64_t mask = half >= 32 ? 0xFFFFFFFFull : ((1ull << half) - 1);
  uint64_t l = (x >> half) & mask, r = x & mask;
  for (int round = 0; round < 4; ++round) {
    uint64_t f = mix64(r ^ (seed + 0xA24BAED4963EE407ull * (uint64_t)(round + 1))) & mask;
    uint64_t nl = r;
    r = l ^ f;
    l = nl;
  }
  return (l << half) | r;
}

__global__ void distinct_keys_kernel(int64_t *keys, uint64_t n, uint64_t domain,
                                     int bits, uint64_t seed, uint64_t offset) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = offset + t;
    if (domain <= 1) {
      keys[t] = 0;
      continue;
    }
    do {
      x = feistel(x, bits, seed);
    } while (x >= domain);
    keys[t] = (int64_t)x;
  }
}

__global__ void uniform_kernel(double *vals, uint64_t n, uint64_t seed, uint64_t offset) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t m = rand_u64(seed, offset + t) >> 11;
    vals[t] = (double)m * 0x1p-52 - 1.0;
  }
}

// Instance i owns entries [off[i], off[i+1]): keys permute_index(t, dom[i],
// kseed + i) and values uniform_pm1(vseed + i, t) for t = 0, 1, ...
__global__ void batch_gen_kernel(int64_t *keys, double *vals, const uint64_t *off, const uint64_t *dom,
                                 uint32_t count, uint64_t n, uint64_t kseed, uint64_t vseed) {
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo = 0, hi = count;  // largest i with off[i] <= p
    while (hi - lo > 1) {
      uint64_t mid = (lo + hi) >> 1;
      if (off[mid] <= p) lo = mid;
      else hi = mid;
    }
    const uint64_t t = p - off[lo], domain = dom[lo];
    if (keys) {
      int bits = 0;
      while (bits < 64 && (1ull << bits) < domain) ++bits;
      if (bits & 1) ++bits;
      if (bits < 2) bits = 2;
      uint64_t x = t;
      if (domain <= 1) {
        x = 0;
      } else {
        do {
          x = feistel(x, bits, kseed + lo);
        } while (x >= domain);
      }
      keys[p] = (int64_t)x;
    }
    if (vals) vals[p] = (double)(rand_u64(vseed + lo, t) >> 11) * 0x1p-52 - 1.0;
  }
}

}  // namespace
}  // namespace lnt

using namespace lnt;

extern "C" int lnt_testgen_batch(lnt_context ctx, int64_t *keys, double *vals, const int64_t *offsets,
                                 const uint64_t *domains, int64_t count, uint64_t key_seed, uint64_t val_seed) {
  if (!ctx || count < 0 || (count > 0 && (!offsets || !domains)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "bad testgen arguments");
  if (count == 0) return LNT_OK;
  if (offsets[0] != 0) return fail(LNT_ERR_INVALID_ARGUMENT, "offsets[0] must be 0");
  for (int64_t i = 0; i < count; ++i)
    if (offsets[i + 1] < offsets[i] || domains[i] == 0 || (uint64_t)(offsets[i + 1] - offsets[i]) > domains[i])
      return fail(LNT_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": entries exceed the domain");
  const uint64_t n = (uint64_t)offsets[count];
  if (n == 0) return LNT_OK;
  if ((!keys && !vals)) return fail(LNT_ERR_INVALID_ARGUMENT, "no output array");
  LNT_HIP(hipSetDevice(ctx->device));
  void *tab = nullptr;
  const size_t bytes = (2 * (size_t)count + 1) * sizeof(uint64_t);
  LNT_HIP(hipMalloc(&tab, bytes));
  uint64_t *off = (uint64_t *)tab, *dom = off + count + 1;
  int rc = LNT_OK;
  if (hipMemcpy(off, offsets, (count + 1) * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dom, domains, count * 8, hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(LNT_ERR_DEVICE, "hipMemcpy of the generator table failed");
  } else {
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    batch_gen_kernel<<<(uint32_t)blocks, 256, 0, ctx->stream>>>(keys, vals, off, dom, (uint32_t)count, n, key_seed,
                                                                 val_seed);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(LNT_ERR_DEVICE, "batch generator launch failed");
  }
  (void)hipFree(tab);
  return rc;
}

extern "C" int lnt_testgen_distinct_keys(lnt_context ctx, int64_t *keys, int64_t n,
                                         uint64_t domain, uint64_t seed, uint64_t offset) {
  if (!ctx || !keys || n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "bad testgen arguments");
  if (domain == 0 || offset + (uint64_t)n > domain)
    return fail(LNT_ERR_INVALID_ARGUMENT, "offset + n exceeds the domain");
  if (n == 0) return LNT_OK;
  int bits = 0;
  while (bits < 64 && (1ull << bits) < domain) ++bits;
  if (bits & 1) ++bits;
  if (bits < 2) bits = 2;
  LNT_HIP(hipSetDevice(ctx->device));
  uint64_t blocks = ((uint64_t)n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  distinct_keys_kernel<<<(uint32_t)blocks, 256, 0, ctx->stream>>>(keys, (uint64_t)n, domain, bits, seed, offset);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

extern "C" int lnt_testgen_uniform(lnt_context ctx, double *vals, int64_t n, uint64_t seed,
                                   uint64_t offset) {
  if (!ctx || !vals || n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "bad testgen arguments");
  if (n == 0) return LNT_OK;
  LNT_HIP(hipSetDevice(ctx->device));
  uint64_t blocks = ((uint64_t)n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  uniform_kernel<<<(uint32_t)blocks, 256, 0, ctx->stream>>>(vals, (uint64_t)n, seed, offset);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}
