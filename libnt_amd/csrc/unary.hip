// unary.hip -- unary contraction and attraction of one sparse operand
// (lnt_contract_attract): B(j,k) = A(i,j,i,k) sums the diagonal of the
// repeated index, B(j,i) = A(!i,j,!i) keeps it as one mode.
//
// Replaces SparseMIABase::contract_attract (multi_index_array/
// SparseMIABase.h:1913-2053), which sorts the operand with the contracted
// modes fastest, then the attracted ones, then the rest, and walks each run of
// entries that share the attracted and remaining modes, summing from 0 the
// entries whose contracted indices agree and keeping non-zero sums.
//
// Here: one map kernel decodes every key, drops the entries off the diagonals
// (they add nothing) and gives the others a compressed key -- contraction
// diagonal values lowest, then the output key (remaining modes, then
// attracted ones) -- whose ascending order is the reference's order within
// each output entry; the onesweep sort orders them; one thread per output
// entry sums its run in that order (the same additions as the reference, so
// the same bits); a scan and a compaction write C sorted by key.
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

uint32_t grid_for(uint64_t n) { return (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock, 16384); }

// Per-call description, read by every lane with uniform addresses.
struct UnaryDev {
  int32_t order, ngroups;
  uint64_t sentinel;  // P * O * A: the key of an entry off a diagonal
  uint64_t P, O;      // contraction and remaining dimensionalities
  Divider p_div;
  // per input position t (mode seq[t]): digit divider, group, place value
  // among the remaining modes
  uint64_t magic[LNT_MAX_ORDER], d[LNT_MAX_ORDER];
  uint32_t s1[LNT_MAX_ORDER], s2[LNT_MAX_ORDER];
  int32_t grp[LNT_MAX_ORDER];
  uint64_t oplace[LNT_MAX_ORDER];
  // per group g-1: kind and place value of its diagonal value
  int32_t gkind[LNT_MAX_ORDER];
  uint64_t gplace[LNT_MAX_ORDER];
};

__global__ void unary_map_kernel(const uint64_t *__restrict__ keys, const double *__restrict__ vals, uint32_t n,
                                 const UnaryDev *__restrict__ u, uint64_t *__restrict__ ok,
                                 double *__restrict__ ov) {
  const int order = u->order, ngroups = u->ngroups;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t key = keys[i];
    uint64_t diag[LNT_MAX_ORDER];
    for (int g = 0; g < ngroups; ++g) diag[g] = ~0ull;
    bool on = true;
    uint64_t other = 0;
    for (int t = 0; t < order; ++t) {
      const uint64_t hi = mulhi64(u->magic[t], key);
      const uint64_t q = (hi + ((key - hi) >> u->s1[t])) >> u->s2[t];
      const uint64_t s = key - q * u->d[t];
      key = q;
      const int g = u->grp[t];
      if (g == 0) {
        other += s * u->oplace[t];
      } else if (diag[g - 1] == ~0ull) {
        diag[g - 1] = s;
      } else {
        on = on && diag[g - 1] == s;
      }
    }
    uint64_t con = 0, att = 0;
    for (int g = 0; g < ngroups; ++g) {
      if (u->gkind[g] == LNT_UNARY_CONTRACT) con += diag[g] * u->gplace[g];
      else att += diag[g] * u->gplace[g];
    }
    ok[i] = on ? con + u->P * (other + u->O * att) : u->sentinel;
    ov[i] = vals[i];
  }
}

// One thread per run of keys with one output key (key / P): the run's sum
// from 0 in sorted order, kept when non-zero (SparseMIABase.h:2024-2044).
// The sum overwrites the run's first value; runs are disjoint.
__global__ void unary_sum_kernel(const uint64_t *__restrict__ k, double *__restrict__ v, uint32_t n,
                                 const UnaryDev *__restrict__ u, uint32_t *__restrict__ keep) {
  const uint64_t sentinel = u->sentinel;
  const Divider pd = u->p_div;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = k[p];
    const uint64_t g = divide(key, pd);
    if (key == sentinel || (p > 0 && divide(k[p - 1], pd) == g)) {
      keep[p] = 0;
      continue;
    }
    double sum = 0;
    for (uint64_t q = p; q < n && k[q] != sentinel && divide(k[q], pd) == g; ++q) sum += v[q];
    v[p] = sum;
    keep[p] = sum != 0;
  }
}

__global__ void unary_compact_kernel(const uint64_t *__restrict__ k, const double *__restrict__ v,
                                     const uint32_t *__restrict__ keep, const uint64_t *__restrict__ pos, uint32_t n,
                                     const UnaryDev *__restrict__ u, uint64_t *__restrict__ ck,
                                     double *__restrict__ cv) {
  const Divider pd = u->p_div;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
    if (!keep[p]) continue;
    const uint64_t d = pos[p];
    ck[d] = divide(k[p], pd);
    cv[d] = v[p];
  }
}

}  // namespace
}  // namespace lnt

using namespace lnt;

extern "C" int lnt_contract_attract(lnt_context ctx, const int64_t *keys, const double *vals, int64_t n, int order,
                                    const int64_t *dims, const int32_t *seq, const int32_t *group, const int32_t *kind,
                                    int ngroups, int64_t **c_keys, double **c_vals, int64_t *nc, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!c_keys || !c_vals || !nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  if (n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (n >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "n must be < 2^32");
  if (n && (!keys || !vals)) return fail(LNT_ERR_INVALID_ARGUMENT, "null keys or values");
  if (!dims || !seq || !group || (ngroups && !kind)) return fail(LNT_ERR_INVALID_ARGUMENT, "null descriptor");
  Relin r;
  uint64_t dimensionality;
  LNT_TRY(make_relin(order, dims, seq, seq, &r, &dimensionality));  // validates order, dims, seq
  if (ngroups < 0 || ngroups > order) return fail(LNT_ERR_INVALID_ARGUMENT, "ngroups must be in [0, order]");
  int64_t gdim[LNT_MAX_ORDER];
  for (int g = 0; g < ngroups; ++g) {
    if (kind[g] != LNT_UNARY_CONTRACT && kind[g] != LNT_UNARY_ATTRACT)
      return fail(LNT_ERR_INVALID_ARGUMENT, "kind must be LNT_UNARY_CONTRACT or LNT_UNARY_ATTRACT");
    gdim[g] = 0;
  }
  for (int m = 0; m < order; ++m) {
    const int g = group[m];
    if (g < 0 || g > ngroups) return fail(LNT_ERR_INVALID_ARGUMENT, "group numbers must be in [0, ngroups]");
    if (g == 0) continue;
    if (gdim[g - 1] && gdim[g - 1] != dims[m])
      return fail(LNT_ERR_DIMENSION_MISMATCH, "the modes of one group differ in size");
    gdim[g - 1] = dims[m];
  }
  UnaryDev u{};
  u.order = order;
  u.ngroups = ngroups;
  uint64_t P = 1, A = 1, O = 1;
  for (int g = 0; g < ngroups; ++g) {
    if (!gdim[g]) return fail(LNT_ERR_INVALID_ARGUMENT, "a group has no mode");
    u.gkind[g] = kind[g];
    if (kind[g] == LNT_UNARY_CONTRACT) {
      u.gplace[g] = P;
      P *= (uint64_t)gdim[g];
    } else {
      u.gplace[g] = A;
      A *= (uint64_t)gdim[g];
    }
  }
  uint64_t oplace_of_mode[LNT_MAX_ORDER];
  for (int m = 0; m < order; ++m)
    if (group[m] == 0) {
      oplace_of_mode[m] = O;
      O *= (uint64_t)dims[m];
    }
  // P*O*A divides prod(dims), so it fits.
  u.P = P;
  u.O = O;
  u.sentinel = P * O * A;
  u.p_div = make_divider(P);
  for (int t = 0; t < order; ++t) {
    const int m = seq[t];
    const Divider dv = make_divider((uint64_t)dims[m]);
    u.magic[t] = dv.magic;
    u.d[t] = dv.d;
    u.s1[t] = dv.s1;
    u.s2[t] = dv.s2;
    u.grp[t] = group[m];
    u.oplace[t] = group[m] == 0 ? oplace_of_mode[m] : 0;
  }
  if (u.sentinel + 1 > (uint64_t(1) << 63)) return fail(LNT_ERR_TOO_LARGE, "the output key space exceeds 2^63");

  // Work arrays: compressed keys and values, keep flags, positions.
  const uint64_t nn = (uint64_t)n;
  Staging st;
  LNT_TRY(ensure(st.b[0], sizeof(UnaryDev)));
  LNT_TRY(ensure(ctx->exec_buf[0], (nn ? nn : 1) * 8));
  LNT_TRY(ensure(ctx->exec_buf[1], (nn ? nn : 1) * 8));
  LNT_TRY(ensure(ctx->exec_buf[2], nn * 4 + 256));
  LNT_TRY(ensure(ctx->exec_buf[3], (nn + 1) * 8 + 256));
  LNT_TRY(ensure(ctx->aux, (scan_workspace_words(nn) + 64) * 8));
  const UnaryDev *ud = (const UnaryDev *)st.b[0].ptr;
  uint64_t *wk = (uint64_t *)ctx->exec_buf[0].ptr;
  double *wv = (double *)ctx->exec_buf[1].ptr;
  uint32_t *keep = (uint32_t *)ctx->exec_buf[2].ptr;
  uint64_t *pos = (uint64_t *)ctx->exec_buf[3].ptr;
  const uint64_t *ik = (const uint64_t *)keys;
  const double *iv = vals;
  if (mem == LNT_MEM_HOST && n) {
    LNT_TRY(ensure(st.b[1], nn * 8));
    LNT_TRY(ensure(st.b[2], nn * 8));
    LNT_HIP(hipMemcpyAsync(st.b[1].ptr, keys, nn * 8, hipMemcpyHostToDevice, ctx->stream));
    LNT_HIP(hipMemcpyAsync(st.b[2].ptr, vals, nn * 8, hipMemcpyHostToDevice, ctx->stream));
    ik = (const uint64_t *)st.b[1].ptr;
    iv = (const double *)st.b[2].ptr;
  }
  uint64_t count = 0;
  if (n) {
    LNT_HIP(hipMemcpyAsync(st.b[0].ptr, &u, sizeof(u), hipMemcpyHostToDevice, ctx->stream));
    unary_map_kernel<<<grid_for(nn), kBlock, 0, ctx->stream>>>(ik, iv, (uint32_t)n, ud, wk, wv);
    LNT_HIP(hipGetLastError());
    LNT_TRY(sort_device(ctx, wk, wv, (uint32_t)n, u.sentinel + 1, nullptr));
    unary_sum_kernel<<<grid_for(nn), kBlock, 0, ctx->stream>>>(wk, wv, (uint32_t)n, ud, keep);
    LNT_HIP(hipGetLastError());
    LNT_TRY(exclusive_scan<uint32_t>(ctx, keep, pos, nn, pos + nn, ctx->aux, 0));
    LNT_HIP(hipMemcpyAsync(ctx->pinned + 8, pos + nn, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipStreamSynchronize(ctx->stream));
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
      rc = ensure(ctx->exec_buf[4], count * 8);
      if (rc == LNT_OK) rc = ensure(ctx->exec_buf[5], count * 8);
      dk = (uint64_t *)ctx->exec_buf[4].ptr;
      dv = (double *)ctx->exec_buf[5].ptr;
    }
    if (rc == LNT_OK) {
      unary_compact_kernel<<<grid_for(nn), kBlock, 0, ctx->stream>>>(wk, wv, keep, pos, (uint32_t)n, ud, dk, dv);
      if (hipGetLastError() != hipSuccess) rc = fail(LNT_ERR_DEVICE, "unary_compact_kernel launch failed");
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
