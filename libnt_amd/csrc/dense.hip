// dense.hip -- product of a sparse operand with a dense one when no index is
// contracted (lnt_sparse_dense_mult): C(l, r, t) = A(l, t) * B(r, t) over the
// outer indices l of A, r of B and the shared element-wise indices t.
//
// Replaces SparseMIABase::noLatticeMult with a DenseMIA operand
// (multi_index_array/SparseMIABase.h:1829-1910): for every A entry with a
// non-zero value, in A's order, and every B outer index r in order whose B
// value is non-zero, one C entry a*b at key r + l*R + t*L*R.
//
// Here: B's non-zeros are compacted per inter index t (flags, scan,
// compaction: the list of (r, b) of every t in r order); each A entry then
// owns cnt[t] consecutive outputs (scan over A); one pass writes C in output
// order, kChunk outputs per workgroup round, each output finding its A entry
// by a search over the offsets staged in LDS -- consecutive lanes write
// consecutive entries.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "lnt_internal.h"

namespace lnt {

template <typename T>
int exclusive_scan(lnt_context ctx, const T *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                   DeviceBuffer &tmp, size_t tmp_offset);
size_t scan_workspace_words(uint64_t n);

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kChunk = 2048;

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 16384)); }

// B's layout: offset of (r, t) = sum over B's outer modes of digit(r) * place
// + sum over its inter modes of digit(t) * place (first listed mode fastest).
struct DenseDev {
  int32_t nr, ni;
  uint64_t R, T, L;
  Divider l_div;
  uint64_t r_dim[LNT_MAX_ORDER], r_place[LNT_MAX_ORDER];
  uint64_t i_dim[LNT_MAX_ORDER], i_place[LNT_MAX_ORDER];
};

__device__ inline uint64_t b_offset(const DenseDev *__restrict__ u, uint64_t r, uint64_t t) {
  uint64_t off = 0;
  for (int q = 0; q < u->nr; ++q) {
    const uint64_t d = u->r_dim[q];
    off += (r % d) * u->r_place[q];
    r /= d;
  }
  for (int q = 0; q < u->ni; ++q) {
    const uint64_t d = u->i_dim[q];
    off += (t % d) * u->i_place[q];
    t /= d;
  }
  return off;
}

// flag[t*R + r] = B(r, t) != 0
__global__ void dense_flags_kernel(const double *__restrict__ b, const DenseDev *__restrict__ u,
                                   uint32_t *__restrict__ flag) {
  const uint64_t R = u->R, n = R * u->T;
  for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x)
    flag[x] = b[b_offset(u, x % R, x / R)] != 0.0;
}

// The non-zeros of B in (t, r) order: nz_r / nz_v.
__global__ void dense_compact_kernel(const double *__restrict__ b, const DenseDev *__restrict__ u,
                                     const uint32_t *__restrict__ flag, const uint64_t *__restrict__ pos,
                                     uint64_t *__restrict__ nz_r, double *__restrict__ nz_v) {
  const uint64_t R = u->R, n = R * u->T;
  for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
    if (!flag[x]) continue;
    const uint64_t r = x % R;
    nz_r[pos[x]] = r;
    nz_v[pos[x]] = b[b_offset(u, r, x / R)];
  }
}

// Outputs of each A entry: the non-zeros of B at its inter index, none for a
// zero value (:1881).  A key whose inter index is out of range raises `err`.
__global__ void dense_counts_kernel(const uint64_t *__restrict__ ak, const double *__restrict__ av, uint32_t na,
                                    const DenseDev *__restrict__ u, const uint64_t *__restrict__ pos,
                                    uint64_t *__restrict__ cnt, uint32_t *__restrict__ err) {
  const uint64_t R = u->R, T = u->T;
  const Divider ld = u->l_div;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < na; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = divide(ak[e], ld);
    uint64_t c = 0;
    if (t >= T) atomicOr(err, 1u);
    else if (av[e] != 0.0) c = pos[(t + 1) * R] - pos[t * R];
    cnt[e] = c;
  }
}

// Largest s in [lo, hi) with off[s] <= q (off ascending, off[lo] <= q).
__device__ inline uint64_t upper_index(const uint64_t *__restrict__ off, uint64_t lo, uint64_t hi, uint64_t q) {
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (off[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void dense_write_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, uint32_t na, const DenseDev *__restrict__ u,
    const uint64_t *__restrict__ pos, const uint64_t *__restrict__ nz_r, const double *__restrict__ nz_v,
    const uint64_t *__restrict__ out_off, uint64_t nc, uint64_t *__restrict__ ck, double *__restrict__ cv) {
  __shared__ uint64_t s_off[kChunk + 1];
  __shared__ uint64_t s_rng[2];
  const uint64_t R = u->R, L = u->L;
  const Divider ld = u->l_div;
  for (uint64_t q0 = (uint64_t)blockIdx.x * kChunk; q0 < nc; q0 += (uint64_t)gridDim.x * kChunk) {
    const uint64_t q1 = min<uint64_t>(q0 + kChunk, nc);
    if (threadIdx.x == 0) {
      const uint64_t e0 = upper_index(out_off, 0, na, q0);
      s_rng[0] = e0;
      s_rng[1] = upper_index(out_off, e0, na, q1 - 1) + 1;
    }
    __syncthreads();
    const uint64_t e0 = s_rng[0], e1 = s_rng[1];
    const bool staged = e1 - e0 <= kChunk + 1;
    if (staged)
      for (uint64_t x = threadIdx.x; x < e1 - e0; x += blockDim.x) s_off[x] = out_off[e0 + x];
    __syncthreads();
    for (uint64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
      uint64_t e, base;
      if (staged) {
        uint32_t lo = 0, hi = (uint32_t)(e1 - e0);
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_off[mid] <= q) lo = mid;
          else hi = mid;
        }
        e = e0 + lo;
        base = s_off[lo];
      } else {
        e = upper_index(out_off, e0, e1, q);
        base = out_off[e];
      }
      const uint64_t key = ak[e];
      const uint64_t t = divide(key, ld), l = key - t * L;
      const uint64_t p = pos[t * R] + (q - base);
      ck[q] = nz_r[p] + l * R + t * L * R;
      cv[q] = av[e] * nz_v[p];
    }
    __syncthreads();  // s_off is restaged next round
  }
}

}  // namespace
}  // namespace lnt

using namespace lnt;

extern "C" int lnt_sparse_dense_mult(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                                     int64_t l_size, int order_b, const int64_t *b_dims, const double *b,
                                     const int32_t *r_modes, int nr, const int32_t *i_modes, int ni,
                                     int64_t **c_keys, double **c_vals, int64_t *nc, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!c_keys || !c_vals || !nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  if (na < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "na must be < 2^32");
  if (na && (!a_keys || !a_vals)) return fail(LNT_ERR_INVALID_ARGUMENT, "null A arrays");
  if (l_size < 1) return fail(LNT_ERR_INVALID_ARGUMENT, "l_size must be >= 1");
  if (order_b < 1 || order_b > LNT_MAX_ORDER || !b_dims || !b)
    return fail(LNT_ERR_INVALID_ARGUMENT, "B needs 1..LNT_MAX_ORDER modes, dims and data");
  if (nr < 0 || ni < 0 || nr + ni != order_b || (nr && !r_modes) || (ni && !i_modes))
    return fail(LNT_ERR_INVALID_ARGUMENT, "B's outer and inter modes must list each of its modes once");
  bool seen[LNT_MAX_ORDER] = {};
  uint64_t place[LNT_MAX_ORDER];
  unsigned __int128 size = 1;
  for (int m = 0; m < order_b; ++m) {
    if (b_dims[m] < 1) return fail(LNT_ERR_INVALID_ARGUMENT, "B dims must be >= 1");
    place[m] = (uint64_t)size;
    size *= (uint64_t)b_dims[m];
    if (size >= ((unsigned __int128)1 << 40)) return fail(LNT_ERR_TOO_LARGE, "dense B holds at most 2^40 elements");
  }
  DenseDev u{};
  u.nr = nr;
  u.ni = ni;
  u.R = u.T = 1;
  for (int q = 0; q < nr + ni; ++q) {
    const int m = q < nr ? r_modes[q] : i_modes[q - nr];
    if (m < 0 || m >= order_b || seen[m])
      return fail(LNT_ERR_INVALID_ARGUMENT, "B's outer and inter modes must list each of its modes once");
    seen[m] = true;
    if (q < nr) {
      u.r_dim[q] = (uint64_t)b_dims[m];
      u.r_place[q] = place[m];
      u.R *= (uint64_t)b_dims[m];
    } else {
      u.i_dim[q - nr] = (uint64_t)b_dims[m];
      u.i_place[q - nr] = place[m];
      u.T *= (uint64_t)b_dims[m];
    }
  }
  u.L = (uint64_t)l_size;
  u.l_div = make_divider(u.L);
  if ((unsigned __int128)u.L * u.R * u.T >= ((unsigned __int128)1 << 63))
    return fail(LNT_ERR_TOO_LARGE, "C's key space exceeds 2^63");
  const uint64_t nbd = u.R * u.T;  // B elements
  const uint64_t nn = (uint64_t)na;

  Staging st;
  LNT_TRY(ensure(st.b[0], sizeof(DenseDev)));
  LNT_TRY(ensure(st.b[1], nbd * 4 + 256));                  // flags
  LNT_TRY(ensure(st.b[2], (nbd + 1) * 8 + 256));            // positions of B's non-zeros
  LNT_TRY(ensure(st.b[3], (nn + 1) * 8 + 256));             // A's output counts, then offsets
  LNT_TRY(ensure(ctx->aux, (scan_workspace_words(std::max(nbd, nn)) + 64) * 8));
  LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
  DenseDev *ud = (DenseDev *)st.b[0].ptr;
  uint32_t *flag = (uint32_t *)st.b[1].ptr;
  uint64_t *pos = (uint64_t *)st.b[2].ptr;
  uint64_t *off = (uint64_t *)st.b[3].ptr;
  uint32_t *err = (uint32_t *)ctx->small.ptr + kSmallFlags + 6;
  LNT_HIP(hipMemcpyAsync(ud, &u, sizeof(u), hipMemcpyHostToDevice, ctx->stream));
  const uint64_t *ak = (const uint64_t *)a_keys;
  const double *av = a_vals, *bd = b;
  if (mem == LNT_MEM_HOST) {
    LNT_TRY(ensure(st.b[4], (nn ? nn : 1) * 16));
    LNT_TRY(ensure(st.b[5], nbd * 8));
    if (nn) {
      LNT_HIP(hipMemcpyAsync(st.b[4].ptr, a_keys, nn * 8, hipMemcpyHostToDevice, ctx->stream));
      LNT_HIP(hipMemcpyAsync((char *)st.b[4].ptr + nn * 8, a_vals, nn * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    LNT_HIP(hipMemcpyAsync(st.b[5].ptr, b, nbd * 8, hipMemcpyHostToDevice, ctx->stream));
    ak = (const uint64_t *)st.b[4].ptr;
    av = (const double *)((char *)st.b[4].ptr + nn * 8);
    bd = (const double *)st.b[5].ptr;
  }
  // B's non-zeros per inter index; pos[t*R] is where index t's list starts,
  // pos[R*T] the total.
  dense_flags_kernel<<<grid_for(nbd), kBlock, 0, ctx->stream>>>(bd, ud, flag);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, flag, pos, nbd, pos + nbd, ctx->aux, 0));
  uint64_t nnz_b = 0, count = 0;
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 8, pos + nbd, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(&nnz_b, ctx->pinned + 8, sizeof(nnz_b));
  LNT_TRY(ensure(ctx->exec_buf[4], (nnz_b ? nnz_b : 1) * 8));
  LNT_TRY(ensure(ctx->exec_buf[5], (nnz_b ? nnz_b : 1) * 8));
  uint64_t *nz_r = (uint64_t *)ctx->exec_buf[4].ptr;
  double *nz_v = (double *)ctx->exec_buf[5].ptr;
  if (nnz_b) {
    dense_compact_kernel<<<grid_for(nbd), kBlock, 0, ctx->stream>>>(bd, ud, flag, pos, nz_r, nz_v);
    LNT_HIP(hipGetLastError());
  }
  if (nn) {
    LNT_TRY(launch_fill_u32(ctx, err, 0, 1));
    dense_counts_kernel<<<grid_for(nn), kBlock, 0, ctx->stream>>>(ak, av, (uint32_t)na, ud, pos, off, err);
    LNT_HIP(hipGetLastError());
    LNT_TRY(exclusive_scan<uint64_t>(ctx, off, off, nn, off + nn, ctx->aux, 0));
    LNT_HIP(hipMemcpyAsync(ctx->pinned + 8, off + nn, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipMemcpyAsync(ctx->pinned + 10, err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->pinned[10]) return fail(LNT_ERR_KEY_RANGE, "an A key lies beyond l_size * prod(inter dims)");
    memcpy(&count, ctx->pinned + 8, sizeof(count));
  }
  const size_t bytes = (size_t)(count ? count : 1) * 8;
  void *ck = nullptr, *cv = nullptr;
  if (mem == LNT_MEM_DEVICE) {
    if (hipMalloc(&ck, bytes) != hipSuccess || hipMalloc(&cv, bytes) != hipSuccess) {
      (void)hipGetLastError();
      if (ck) (void)hipFree(ck);
      return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of the result failed");
    }
  } else {
    ck = malloc(bytes);
    cv = malloc(bytes);
    if (!ck || !cv) {
      free(ck);
      free(cv);
      return fail(LNT_ERR_OUT_OF_MEMORY, "malloc of the result failed");
    }
  }
  int rc = LNT_OK;
  if (count) {
    uint64_t *dk = (uint64_t *)ck;
    double *dv = (double *)cv;
    if (mem == LNT_MEM_HOST) {
      rc = ensure(ctx->spg_tmp, count * 16);
      dk = (uint64_t *)ctx->spg_tmp.ptr;
      dv = (double *)(dk + count);
    }
    if (rc == LNT_OK) {
      const uint32_t grid = (uint32_t)std::min<uint64_t>((count + kChunk - 1) / kChunk, 16384);
      dense_write_kernel<<<grid, kBlock, 0, ctx->stream>>>(ak, av, (uint32_t)na, ud, pos, nz_r, nz_v, off, count,
                                                            dk, dv);
      if (hipGetLastError() != hipSuccess) rc = fail(LNT_ERR_DEVICE, "dense_write_kernel launch failed");
    }
    if (rc == LNT_OK && mem == LNT_MEM_HOST &&
        (hipMemcpyAsync(ck, dk, count * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
         hipMemcpyAsync(cv, dv, count * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess))
      rc = fail(LNT_ERR_DEVICE, "copy of the result to the host failed");
    if (rc == LNT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(LNT_ERR_DEVICE, "hipStreamSynchronize failed");
  }
  if (rc != LNT_OK) {
    if (mem == LNT_MEM_DEVICE) {
      (void)hipFree(ck);
      (void)hipFree(cv);
    } else {
      free(ck);
      free(cv);
    }
    return rc;
  }
  *c_keys = (int64_t *)ck;
  *c_vals = (double *)cv;
  *nc = (int64_t)count;
  return LNT_OK;
}
