// mex.hip -- the reference's five mex entry points (include/lnt_mex.h) over
// the library's device paths.  The arguments are a mexFunction's host arrays:
// they are staged through device memory, the work runs in the HIP kernels of
// sort.hip / merge.hip / spgemm.hip, and results come back into the caller's
// arrays (in-place entries) or into arrays from the caller's allocator (entries
// whose mex returns new mxArrays).  There is no CPU fallback.
#include <math.h>
#include <stdlib.h>

#include <string>

#include "lnt_internal.h"
#include "../../include/lnt_mex.h"

namespace lnt {
namespace {

constexpr int kBlock = 256;

uint32_t grid_for(uint32_t n) { return std::min<uint32_t>((n + kBlock - 1) / kBlock, 4096); }

// Largest key, read as unsigned: a negative key reads as >= 2^63.
__global__ void key_max_kernel(const uint64_t *__restrict__ keys, uint32_t n, unsigned long long *out) {
  uint64_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    m = max(m, keys[i]);
  for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, (uint64_t)__shfl_xor((unsigned long long)m, o));
  if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicMax(out, (unsigned long long)m);
}

// Sets *flag when two neighbouring keys of a sorted array are equal.
__global__ void has_duplicates_kernel(const uint64_t *__restrict__ keys, uint32_t n, uint32_t *flag) {
  bool dup = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n; i += (uint64_t)gridDim.x * blockDim.x)
    dup |= keys[i] == keys[i + 1];
  if (__ballot(dup) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flag, 1u);
}

// ------------------------------------------------- the reference's Introsort
// Introsort with a follower, as SparseSortMex runs it (sparse_sort.cpp:14-20 ->
// LibMIAAlgorithm.h:752-764): IntrosortRec (:570-611; partition :439-511,
// median of three :547-560, heap sort :522-535, depth 2*floor(log2 n)+2 --
// the bit length times two, :621-632) and a final stable insertion pass
// (:641-660).  Equal keys end where the reference's comparisons and swaps put
// them, which no other sort reproduces; the library runs it only when the
// input holds equal keys (unique keys have one sorted order).  One thread does
// every step in the reference's order, on a copy in LDS when the pairs fit.
struct Pairs {
  uint64_t *k;
  double *v;
  __device__ void swap(int64_t i, int64_t j) const {
    const double tv = v[i];
    v[i] = v[j];
    v[j] = tv;
    const uint64_t tk = k[i];
    k[i] = k[j];
    k[j] = tk;
  }
};

__device__ void intro_sift_down(const Pairs &p, int64_t base, int64_t size, int64_t root) {
  while (root * 2 + 1 < size) {
    int64_t child = root * 2 + 1;
    if (child + 1 < size && p.k[base + child] < p.k[base + child + 1]) child += 1;
    if (p.k[base + root] < p.k[base + child]) {
      p.swap(base + child, base + root);
      root = child;
    } else {
      return;
    }
  }
}

__device__ void intro_heap_sort(const Pairs &p, int64_t base, int64_t size) {
  for (int64_t s = (size - 2) / 2; s >= 0; --s) intro_sift_down(p, base, size, s);
  for (int64_t e = size - 1; e > 0; --e) {
    p.swap(base + e, base);
    intro_sift_down(p, base, e, 0);
  }
}

__device__ int64_t intro_median(const Pairs &p, int64_t one, int64_t two, int64_t three) {
  const bool c12 = p.k[one] < p.k[two], c13 = p.k[one] < p.k[three], c23 = p.k[two] < p.k[three];
  if (c12 && c23) return two;
  if (c12 && !c23 && c13) return three;
  if (!c12 && c13) return one;
  if (!c12 && !c13 && c23) return three;
  if (c12 && !c13) return one;
  return two;
}

__device__ int64_t intro_partition(const Pairs &p, int64_t b, int64_t e) {
  int64_t lhs = b + 1, rhs = e - 1;
  for (;;) {
    while (lhs < rhs && !(p.k[rhs] < p.k[b])) --rhs;
    while (lhs < rhs && p.k[lhs] < p.k[b]) ++lhs;
    if (lhs == rhs) break;
    p.swap(lhs, rhs);
  }
  if (!(p.k[lhs] < p.k[b])) return b;
  p.swap(b, lhs);
  return lhs;
}

// The recursion's subarrays are disjoint, so an explicit stack visits them
// in any order with the same result; each keeps its own depth budget.
__device__ void intro_sort(const Pairs &p, int64_t n) {
  struct Range {
    int64_t b, e;
    uint32_t depth;
  } stack[80];
  int top = 0;
  uint32_t depth = 0;
  for (uint64_t m = (uint64_t)n; m; m >>= 1) ++depth;
  stack[top++] = {0, n, depth * 2};
  while (top) {
    Range r = stack[--top];
    for (;;) {
      const int64_t len = r.e - r.b;
      if (len < 50) break;
      if (r.depth == 0) {
        intro_heap_sort(p, r.b, len);
        break;
      }
      const int64_t pivot = intro_median(p, r.b, r.b + len / 2, r.e - 1);
      p.swap(pivot, r.b);
      const int64_t q = intro_partition(p, r.b, r.e);
      stack[top++] = {q + 1, r.e, r.depth - 1};
      r = {r.b, q, r.depth - 1};
    }
  }
  for (int64_t i = 1; i < n; ++i)
    for (int64_t t = i; t > 0 && p.k[t] < p.k[t - 1]; --t) p.swap(t, t - 1);
}

constexpr uint32_t kIntroLds = 8192;       // pairs sorted in LDS (128 KiB)
constexpr uint32_t kIntroMax = 1u << 16;   // largest input replayed exactly

__global__ __launch_bounds__(kBlock) void introsort_kernel(uint64_t *__restrict__ keys, double *__restrict__ vals,
                                                           uint32_t n) {
  __shared__ uint64_t s_k[kIntroLds];
  __shared__ double s_v[kIntroLds];
  if (n <= kIntroLds) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      s_k[i] = keys[i];
      s_v[i] = vals[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) intro_sort(Pairs{s_k, s_v}, n);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      keys[i] = s_k[i];
      vals[i] = s_v[i];
    }
  } else if (threadIdx.x == 0) {
    intro_sort(Pairs{keys, vals}, n);
  }
}

// v[i] = -1 * v[i]: the subtraction from an empty left operand
// (MATLAB/mex/sparse_merge.cpp:52-54).
__global__ void negate_kernel(double *v, uint32_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    v[i] = -1.0 * v[i];
}

// 1-based uint8 mode sequences of the mex side, as 0-based int32 (0 becomes
// -1, which the permutation check rejects).
void zero_based(const uint8_t *s, int order, int32_t *out) {
  for (int t = 0; t < order; ++t) out[t] = (int32_t)s[t] - 1;
}

bool same_seq(const uint8_t *a, const uint8_t *b, int order) {
  for (int t = 0; t < order; ++t)
    if (a[t] != b[t]) return false;
  return true;
}

int check_seq_args(int64_t n, const uint8_t *cur, const uint8_t *nxt, const int64_t *dims, int order) {
  if (n >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "n must be < 2^32");
  if (order < 1 || order > LNT_MAX_ORDER)
    return fail(LNT_ERR_INVALID_ARGUMENT, "order must be in [1, " + std::to_string(LNT_MAX_ORDER) + "]");
  if (!cur || !nxt || !dims) return fail(LNT_ERR_INVALID_ARGUMENT, "null sequence or dims");
  return LNT_OK;
}

// Copies n pairs between the caller's host arrays and staged device arrays.
int stage_in(lnt_context ctx, Staging &st, int slot, const int64_t *k, const double *v, int64_t n) {
  if (n == 0) return LNT_OK;
  const size_t bytes = (size_t)n * 8;
  LNT_TRY(ensure(st.b[slot], bytes));
  LNT_TRY(ensure(st.b[slot + 1], bytes));
  LNT_HIP(hipMemcpyAsync(st.b[slot].ptr, k, bytes, hipMemcpyHostToDevice, ctx->stream));
  LNT_HIP(hipMemcpyAsync(st.b[slot + 1].ptr, v, bytes, hipMemcpyHostToDevice, ctx->stream));
  return LNT_OK;
}

int copy_out(lnt_context ctx, int64_t *k, double *v, const void *dk, const void *dv, int64_t n) {
  const size_t bytes = (size_t)n * 8;
  LNT_HIP(hipMemcpyAsync(k, dk, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipMemcpyAsync(v, dv, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  return LNT_OK;
}

// Hands a device result of n pairs to the caller in arrays from `alloc`
// (calloc when null); an empty result hands back null pointers.
int hand_out(lnt_context ctx, lnt_mex_alloc alloc, const void *dk, const void *dv, int64_t n, double **c_data,
             int64_t **c_idx, int64_t *nc) {
  double *v = nullptr;
  int64_t *k = nullptr;
  if (n > 0) {
    lnt_mex_alloc a = alloc ? alloc : calloc;
    v = (double *)a((size_t)n, sizeof(double));
    k = v ? (int64_t *)a((size_t)n, sizeof(int64_t)) : nullptr;
    if (!v || !k) {
      // Only calloc's blocks can be given back here; a caller-supplied
      // allocator (mxCalloc) reclaims its own.
      if (!alloc) {
        free(v);
        free(k);
      }
      return fail(LNT_ERR_OUT_OF_MEMORY, "allocation of the " + std::to_string(n) + "-element result failed");
    }
    int rc = copy_out(ctx, k, v, dk, dv, n);
    if (rc != LNT_OK) {
      if (!alloc) {
        free(v);
        free(k);
      }
      return rc;
    }
  }
  *c_data = v;
  *c_idx = k;
  *nc = n;
  return LNT_OK;
}

}  // namespace
}  // namespace lnt

using namespace lnt;

extern "C" {

int lnt_mex_sparse_sort(lnt_context ctx, double *data, int64_t *idx, int64_t n) {
  LNT_CTX(ctx);
  if (n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (n >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "n must be < 2^32");
  if (n == 0) return LNT_OK;
  if (!data || !idx) return fail(LNT_ERR_INVALID_ARGUMENT, "null data or indices");
  // The mex takes no dimensions: the sort's key range is the largest index
  // plus one, found on the device.
  Staging st;
  LNT_TRY(stage_in(ctx, st, 0, idx, data, n));
  LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
  // Two flag words (8-byte aligned) of the context's small buffer.
  unsigned long long *mx = (unsigned long long *)((uint32_t *)ctx->small.ptr + kSmallFlags + 4);
  LNT_TRY(launch_fill_u32(ctx, (uint32_t *)mx, 0, 2));
  key_max_kernel<<<grid_for((uint32_t)n), kBlock, 0, ctx->stream>>>((const uint64_t *)st.b[0].ptr, (uint32_t)n, mx);
  LNT_HIP(hipGetLastError());
  uint64_t top = 0;
  LNT_HIP(hipMemcpyAsync(&top, mx, sizeof(top), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  if (top >> 63) return fail(LNT_ERR_KEY_RANGE, "an index is negative");
  LNT_TRY(sort_device(ctx, (uint64_t *)st.b[0].ptr, (double *)st.b[1].ptr, (uint32_t)n, top + 1, nullptr));
  // Equal indices: the radix sort kept their input order, the reference's
  // Introsort leaves them in the order its swaps produce.  Start again from
  // the caller's arrays and replay that Introsort (up to kIntroMax entries:
  // one thread replays it, beyond LDS in global memory).
  uint32_t *dup = (uint32_t *)ctx->small.ptr + kSmallFlags + 3;
  LNT_TRY(launch_fill_u32(ctx, dup, 0, 1));
  has_duplicates_kernel<<<grid_for((uint32_t)n), kBlock, 0, ctx->stream>>>((const uint64_t *)st.b[0].ptr,
                                                                          (uint32_t)n, dup);
  LNT_HIP(hipGetLastError());
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 3, dup, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->pinned[3] && (uint64_t)n <= kIntroMax) {
    LNT_TRY(stage_in(ctx, st, 0, idx, data, n));
    introsort_kernel<<<1, kBlock, 0, ctx->stream>>>((uint64_t *)st.b[0].ptr, (double *)st.b[1].ptr, (uint32_t)n);
    LNT_HIP(hipGetLastError());
  }
  return copy_out(ctx, idx, data, st.b[0].ptr, st.b[1].ptr, n);
}

int lnt_mex_sparse_change_linidx(lnt_context ctx, int64_t *idx, int64_t n, const uint8_t *cur_seq,
                                 const uint8_t *new_seq, const int64_t *dims, int order) {
  LNT_CTX(ctx);
  if (n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (n == 0) return LNT_OK;  // sparse_change_linIdx.cpp:30-31
  LNT_TRY(check_seq_args(n, cur_seq, new_seq, dims, order));
  if (same_seq(cur_seq, new_seq, order)) return LNT_OK;  // :49-50
  int32_t cur[LNT_MAX_ORDER], nxt[LNT_MAX_ORDER];
  zero_based(cur_seq, order, cur);
  zero_based(new_seq, order, nxt);
  return lnt_change_linidx(ctx, idx, n, order, dims, cur, nxt, LNT_MEM_HOST);
}

int lnt_mex_sparse_permute(lnt_context ctx, double *data, int64_t *idx, int64_t n, const uint8_t *cur_seq,
                           const uint8_t *new_seq, const int64_t *dims, int order) {
  LNT_CTX(ctx);
  if (n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (n == 0) return LNT_OK;  // sparse_permute.cpp:57-58
  LNT_TRY(check_seq_args(n, cur_seq, new_seq, dims, order));
  if (same_seq(cur_seq, new_seq, order)) return LNT_OK;  // :76-77
  if (!data || !idx) return fail(LNT_ERR_INVALID_ARGUMENT, "null data or indices");
  int32_t cur[LNT_MAX_ORDER], nxt[LNT_MAX_ORDER];
  zero_based(cur_seq, order, cur);
  zero_based(new_seq, order, nxt);
  Relin r;
  uint64_t dimensionality;
  LNT_TRY(make_relin(order, dims, cur, nxt, &r, &dimensionality));
  // The indices are already in the new linearisation: no relinearisation,
  // and the low bits the old order already sorts need no radix pass.
  const uint32_t low = sorted_low_bits(order, dims, cur, nxt);
  Staging st;
  LNT_TRY(stage_in(ctx, st, 0, idx, data, n));
  LNT_TRY(sort_device(ctx, (uint64_t *)st.b[0].ptr, (double *)st.b[1].ptr, (uint32_t)n, dimensionality, nullptr,
                      low));
  return copy_out(ctx, idx, data, st.b[0].ptr, st.b[1].ptr, n);
}

int lnt_mex_sparse_merge(lnt_context ctx, const double *a_data, const int64_t *a_idx, int64_t na,
                         const double *b_data, const int64_t *b_idx, int64_t nb, double op, lnt_mex_alloc alloc,
                         double **c_data, int64_t **c_idx, int64_t *nc) {
  LNT_CTX(ctx);
  if (!c_data || !c_idx || !nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na + nb >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "na + nb must be < 2^32");
  if ((na && (!a_data || !a_idx)) || (nb && (!b_data || !b_idx)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null operand array");
  const double f = floor(op);  // check_sparse.cpp:121-123
  if (!(f == 0.0 || f == 1.0))
    return fail(LNT_ERR_INVALID_ARGUMENT, "op must be 0 (add) or 1 (subtract)");
  if (na == 0 && nb == 0) return hand_out(ctx, alloc, nullptr, nullptr, 0, c_data, c_idx, nc);
  Staging st;
  if (na == 0 || nb == 0) {
    // One operand empty: the other one is the result, copied (B negated for
    // subtraction), sparse_merge.cpp:29-57.
    const int64_t n = na ? na : nb;
    LNT_TRY(stage_in(ctx, st, 0, na ? a_idx : b_idx, na ? a_data : b_data, n));
    if (na == 0 && f == 1.0) {
      negate_kernel<<<grid_for((uint32_t)n), kBlock, 0, ctx->stream>>>((double *)st.b[1].ptr, (uint32_t)n);
      LNT_HIP(hipGetLastError());
    }
    return hand_out(ctx, alloc, st.b[0].ptr, st.b[1].ptr, n, c_data, c_idx, nc);
  }
  LNT_TRY(stage_in(ctx, st, 0, a_idx, a_data, na));
  LNT_TRY(stage_in(ctx, st, 2, b_idx, b_data, nb));
  LNT_TRY(ensure(st.b[4], (size_t)(na + nb) * 8));
  LNT_TRY(ensure(st.b[5], (size_t)(na + nb) * 8));
  int64_t count = 0;
  LNT_TRY(lnt_merge_into(ctx, (const int64_t *)st.b[0].ptr, (const double *)st.b[1].ptr, na,
                         (const int64_t *)st.b[2].ptr, (const double *)st.b[3].ptr, nb,
                         f == 0.0 ? LNT_OP_ADD : LNT_OP_SUB, (int64_t *)st.b[4].ptr, (double *)st.b[5].ptr, na + nb,
                         &count, LNT_MEM_DEVICE));
  return hand_out(ctx, alloc, st.b[4].ptr, st.b[5].ptr, count, c_data, c_idx, nc);
}

int lnt_mex_sparselattice_mult(lnt_context ctx, const double *a_data, const int64_t *a_idx, int64_t na,
                               const int64_t a_dims[3], const double *b_data, const int64_t *b_idx, int64_t nb,
                               const int64_t b_dims[3], double algorithm, lnt_mex_alloc alloc, double **c_data,
                               int64_t **c_idx, int64_t *nc) {
  LNT_CTX(ctx);
  if (!c_data || !c_idx || !nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na >= (int64_t(1) << 32) || nb >= (int64_t(1) << 32))
    return fail(LNT_ERR_TOO_LARGE, "operands must hold < 2^32 entries");
  if ((na && (!a_data || !a_idx)) || (nb && (!b_data || !b_idx)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null operand array");
  // (int) of the double argument, sparselattice_mul.cpp:110; 0..5 accepted
  // (:24-67).
  if (!(algorithm > -1.0 && algorithm < 6.0))
    return fail(LNT_ERR_INVALID_ARGUMENT, "algorithm must be between 0 and 5");
  // The method each case runs: 0 operator* = csc_no_accum<true> (DCSC), 1
  // csc_times<false> (CSC), 2 csc_times<true> (DCSC), 3 csc_no_accum<false>
  // (CSC), 5 mtt_times.  Case 4 runs outer_times and then falls through into
  // case 5 (:52-58, no break), which rebuilds both lattices from the same
  // arguments as column-major lattices and overwrites the result with
  // mtt_times's: algorithm 4 returns what algorithm 5 returns.
  static const int kStrategy[6] = {LNT_SPGEMM_SORT, LNT_SPGEMM_DENSE | LNT_SPGEMM_CSC, LNT_SPGEMM_DENSE,
                                   LNT_SPGEMM_SORT | LNT_SPGEMM_CSC, LNT_SPGEMM_MTT, LNT_SPGEMM_MTT};
  const int strategy = kStrategy[(int)algorithm];
  Staging st;
  LNT_TRY(stage_in(ctx, st, 0, a_idx, a_data, na));
  LNT_TRY(stage_in(ctx, st, 2, b_idx, b_data, nb));
  int64_t *dk = nullptr;
  double *dv = nullptr;
  int64_t count = 0;
  LNT_TRY(lnt_lattice_mult(ctx, (const int64_t *)st.b[0].ptr, (const double *)st.b[1].ptr, na, a_dims,
                           (const int64_t *)st.b[2].ptr, (const double *)st.b[3].ptr, nb, b_dims, strategy, &dk, &dv,
                           &count, LNT_MEM_DEVICE));
  int rc = hand_out(ctx, alloc, dk, dv, count, c_data, c_idx, nc);
  (void)lnt_free(ctx, dk, LNT_MEM_DEVICE);
  (void)lnt_free(ctx, dv, LNT_MEM_DEVICE);
  return rc;
}

}  // extern "C"
