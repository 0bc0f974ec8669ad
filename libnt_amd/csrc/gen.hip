// gen.hip -- counter-based synthetic-input generators (include/lnt_testgen.h).
// Integer-only and exact-IEEE arithmetic, matching the CPU oracle bit for bit.
#include "../../include/lnt_testgen.h"
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

}  // namespace
}  // namespace lnt

using namespace lnt;

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
  hipSetDevice(ctx->device);
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
  hipSetDevice(ctx->device);
  uint64_t blocks = ((uint64_t)n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  uniform_kernel<<<(uint32_t)blocks, 256, 0, ctx->stream>>>(vals, (uint64_t)n, seed, offset);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}
