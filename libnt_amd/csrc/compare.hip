// compare.hip -- equality of two sorted sparse tensors for gfx950
// (SparseMIABase::compare_with_sparse, SparseMIABase.h:1243-1296).
//
// The reference walks both operands together, skipping entries below the
// sparse tolerance on either side and requiring the remaining keys to pair up
// with the predicate true.  With unique keys that is a condition per key: the
// two sides are both significant or both not (a missing entry counts as not
// significant), and a significant pair satisfies the predicate.  Two passes
// check it: every entry of A looks for its key in B (pairs and A-only keys),
// every entry of B looks for its key in A (B-only keys).  An operand is taken
// in chunks of kChunk entries per workgroup round; the chunk's key range in the
// other operand is found once, and each entry searches only inside it.  Any
// violation sets a flag.
#include <math.h>

#include <algorithm>

#include "lnt_internal.h"

namespace lnt {

namespace {

constexpr uint32_t kChunk = 2048;
constexpr double kSparseTolerance = 1.11e-16;  // SparseTolerance<double>, LibMIAUtil.h:137-150

__device__ inline uint32_t lower_bound_u64(const uint64_t *__restrict__ a, uint32_t lo, uint32_t hi, uint64_t x) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// isEqualFuzzy (FunctionUtil.h:688-698).
__device__ inline bool equal_fuzzy(double a, double b, double prec) {
  if (fabs(a) < 1 || fabs(b) < 1) return fabs(a - b) <= prec;
  return fabs(a - b) <= fmin(fabs(a), fabs(b)) * prec;
}

// pairs = 1: X = A, Y = B; checks A-only keys and every pair.  pairs = 0: X = B,
// Y = A; checks B-only keys.
__global__ __launch_bounds__(256) void compare_side_kernel(const uint64_t *__restrict__ xk,
                                                           const double *__restrict__ xv, uint32_t nx,
                                                           const uint64_t *__restrict__ yk,
                                                           const double *__restrict__ yv, uint32_t ny, int pairs,
                                                           int mode, double prec, uint32_t *__restrict__ flag) {
  __shared__ uint32_t s_y[2];
  for (uint64_t x0 = (uint64_t)blockIdx.x * kChunk; x0 < nx; x0 += (uint64_t)gridDim.x * kChunk) {
    const uint32_t x1 = (uint32_t)min<uint64_t>(x0 + kChunk, nx);
    if (threadIdx.x == 0) {
      const uint32_t y0 = lower_bound_u64(yk, 0, ny, xk[x0]);
      s_y[0] = y0;
      s_y[1] = lower_bound_u64(yk, y0, ny, xk[x1 - 1] + 1);  // keys are below 2^63
    }
    __syncthreads();
    const uint32_t y0 = s_y[0], y1 = s_y[1];
    bool bad = false;
    for (uint32_t x = (uint32_t)x0 + threadIdx.x; x < x1; x += blockDim.x) {
      const uint64_t key = xk[x];
      const double v = xv[x];
      const uint32_t j = lower_bound_u64(yk, y0, y1, key);
      const bool present = j < y1 && yk[j] == key;
      const bool sx = !(fabs(v) <= kSparseTolerance);
      if (!present) {
        bad |= sx;
      } else if (pairs) {
        const double w = yv[j];
        const bool sy = !(fabs(w) <= kSparseTolerance);
        bad |= sx != sy || (sx && !(mode == LNT_CMP_FUZZY ? equal_fuzzy(v, w, prec) : v == w));
      }
    }
    if (bad) atomicOr(flag, 1u);
    __syncthreads();  // s_y is rewritten next round
  }
}

inline uint32_t chunk_grid(uint64_t n) {
  const uint64_t b = (n + kChunk - 1) / kChunk;
  return (uint32_t)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace

int compare_device(lnt_context ctx, const uint64_t *ak, const double *av, uint32_t na, const uint64_t *bk,
                   const double *bv, uint32_t nb, int mode, double prec, int *equal) {
  LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
  uint32_t *flag = (uint32_t *)ctx->small.ptr + kSmallFlags + 3;
  LNT_TRY(launch_fill_u32(ctx, flag, 0, 1));
  if (na) compare_side_kernel<<<chunk_grid(na), 256, 0, ctx->stream>>>(ak, av, na, bk, bv, nb, 1, mode, prec, flag);
  if (nb) compare_side_kernel<<<chunk_grid(nb), 256, 0, ctx->stream>>>(bk, bv, nb, ak, av, na, 0, mode, prec, flag);
  LNT_HIP(hipGetLastError());
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 3, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  *equal = ctx->pinned[3] ? 0 : 1;
  return LNT_OK;
}

}  // namespace lnt
