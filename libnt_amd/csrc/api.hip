// api.hip -- the extern "C" boundary of include/lnt.h: argument validation,
// device/host staging, context and scratch management, error reporting.
// Every compute path below runs the HIP kernels of sort.hip / merge.hip; there
// is no CPU fallback.
#include <stdlib.h>
#include <string.h>

#include <string>

#include "lnt_internal.h"

namespace lnt {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

static thread_local bool g_alloc_forbidden = false;

void set_alloc_forbidden(bool forbidden) { g_alloc_forbidden = forbidden; }

int ensure(DeviceBuffer &b, size_t bytes) {
  if (b.bytes >= bytes) return LNT_OK;
  if (g_alloc_forbidden)
    return fail(LNT_ERR_DEVICE, "internal: workspace of " + std::to_string(bytes) +
                                    " bytes needed while recording a graph (size it before capture)");
  if (b.ptr) {
    hipError_t e = hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return fail(LNT_ERR_DEVICE, std::string("hipFree: ") + hipGetErrorString(e));
  }
  if (bytes == 0) return LNT_OK;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  b.ptr = p;
  b.bytes = bytes;
  return LNT_OK;
}

void release(DeviceBuffer &b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

int make_relin(int order, const int64_t *dims, const int32_t *cur_seq,
               const int32_t *new_seq, Relin *out, uint64_t *dimensionality) {
  if (order < 1 || order > LNT_MAX_ORDER)
    return fail(LNT_ERR_INVALID_ARGUMENT, "order must be in [1, " + std::to_string(LNT_MAX_ORDER) + "]");
  if (!dims || !cur_seq || !new_seq) return fail(LNT_ERR_INVALID_ARGUMENT, "null dims or sequence");
  unsigned __int128 total = 1;
  for (int m = 0; m < order; ++m) {
    if (dims[m] < 1) return fail(LNT_ERR_INVALID_ARGUMENT, "every dimension must be >= 1");
    total *= (uint64_t)dims[m];
    if (total > ((unsigned __int128)1 << 63))
      return fail(LNT_ERR_TOO_LARGE, "dimensionality exceeds 2^63 (int64 linear indices)");
  }
  uint32_t seen_c = 0, seen_n = 0;
  for (int t = 0; t < order; ++t) {
    if (cur_seq[t] < 0 || cur_seq[t] >= order || new_seq[t] < 0 || new_seq[t] >= order)
      return fail(LNT_ERR_INVALID_ARGUMENT, "sequence entry out of range");
    seen_c |= 1u << cur_seq[t];
    seen_n |= 1u << new_seq[t];
  }
  uint32_t all = order == 32 ? ~0u : ((1u << order) - 1);
  if (seen_c != all || seen_n != all)
    return fail(LNT_ERR_INVALID_ARGUMENT, "sequence is not a permutation of the modes");
  Relin r{};
  r.order = order;
  r.active = 0;
  for (int t = 0; t < order; ++t) r.active |= cur_seq[t] != new_seq[t];
  uint64_t stride_of_mode[LNT_MAX_ORDER];
  uint64_t acc = 1;
  for (int t = 0; t < order; ++t) {
    stride_of_mode[new_seq[t]] = acc;
    acc *= (uint64_t)dims[new_seq[t]];
  }
  acc = 1;
  r.pow2 = 1;
  for (int t = 0; t < order; ++t) {
    r.place[t] = make_divider(acc);
    r.weight[t] = stride_of_mode[cur_seq[t]];
    acc *= (uint64_t)dims[cur_seq[t]];
    r.pow2 &= (dims[t] & (dims[t] - 1)) == 0;
  }
  if (r.pow2) {
    uint32_t sh = 0;
    for (int t = 0; t < order; ++t) {
      uint32_t b = 0;
      while ((uint64_t(1) << b) < (uint64_t)dims[cur_seq[t]]) ++b;
      uint32_t ns = 0;
      while ((uint64_t(1) << ns) < r.weight[t]) ++ns;
      r.old_shift[t] = (uint8_t)sh;
      r.bits[t] = (uint8_t)b;
      r.new_shift[t] = (uint8_t)ns;
      sh += b;
    }
  }
  *out = r;
  *dimensionality = (uint64_t)total;
  return LNT_OK;
}

static const char *kProfNames[kProfKinds] = {"histogram", "onesweep", "partition", "merge", "relin",
                                             "lattice_mult"};

static hipEvent_t take_event(lnt_context ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int prof_begin(lnt_context ctx, int kind, double bytes) {
  if (!ctx->profiling || ctx->capturing) return -1;
  ProfRecord r{kind, take_event(ctx), take_event(ctx), bytes};
  if (!r.start || !r.stop || hipEventRecord(r.start, ctx->stream) != hipSuccess) return -1;
  ctx->prof.push_back(r);
  return (int)ctx->prof.size() - 1;
}

void prof_end(lnt_context ctx, int slot) {
  if (slot < 0) return;
  (void)hipEventRecord(ctx->prof[slot].stop, ctx->stream);
}

void prof_set_bytes(lnt_context ctx, int slot, double bytes) {
  if (slot >= 0) ctx->prof[slot].bytes = bytes;
}

static void prof_clear(lnt_context ctx) {
  for (auto &r : ctx->prof) {
    ctx->event_pool.push_back(r.start);
    ctx->event_pool.push_back(r.stop);
  }
  ctx->prof.clear();
}

namespace {

__global__ void check_sorted_kernel(const uint64_t *keys, uint64_t n, uint32_t *err) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    bad |= (int64_t)k < 0;
    if (i + 1 < n) bad |= !(k < keys[i + 1]);
  }
  if (bad) atomicOr(err, 1u);
}

int check_sorted(lnt_context ctx, const uint64_t *a, uint64_t na, const uint64_t *b, uint64_t nb) {
  LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
  uint32_t *err = (uint32_t *)ctx->small.ptr + kSmallFlags + 2;
  LNT_TRY(launch_fill_u32(ctx, err, 0, 1));
  if (na) check_sorted_kernel<<<(uint32_t)std::min<uint64_t>((na + 255) / 256, 4096), 256, 0, ctx->stream>>>(a, na, err);
  if (nb) check_sorted_kernel<<<(uint32_t)std::min<uint64_t>((nb + 255) / 256, 4096), 256, 0, ctx->stream>>>(b, nb, err);
  LNT_HIP(hipGetLastError());
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 2, err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->pinned[2]) return fail(LNT_ERR_NOT_SORTED, "merge input is not strictly ascending");
  return LNT_OK;
}

}  // namespace
}  // namespace lnt

using namespace lnt;

extern "C" {

int lnt_abi_version(void) { return LNT_ABI_VERSION; }

const char *lnt_last_error(void) { return g_last_error.c_str(); }

int lnt_context_create(lnt_context *out, int device, void *stream) {
  if (!out) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return fail(LNT_ERR_NO_DEVICE, "no HIP device is visible");
  }
  if (device < 0 || device >= count) return fail(LNT_ERR_INVALID_ARGUMENT, "device index out of range");
  DeviceGuard guard(device);
  if (!guard.ok) return fail(LNT_ERR_DEVICE, "hipSetDevice failed");
  lnt_context ctx = new (std::nothrow) lnt_context_s();
  if (!ctx) return fail(LNT_ERR_OUT_OF_MEMORY, "context allocation failed");
  ctx->device = device;
  ctx->stream = (hipStream_t)stream;
  if (hipHostMalloc((void **)&ctx->pinned, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
    delete ctx;
    return fail(LNT_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
  }
  memset(ctx->pinned, 0, 64 * sizeof(uint32_t));
  int rc = ensure(ctx->small, kSmallWords * sizeof(uint32_t));
  if (rc == LNT_OK) rc = noalloc_scratch_init(ctx);
  if (rc != LNT_OK) {
    release(ctx->small);
    (void)hipHostFree(ctx->pinned);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return LNT_OK;
}

int lnt_context_destroy(lnt_context ctx) {
  if (!ctx) return LNT_OK;
  DeviceGuard guard(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (DeviceBuffer *b : {&ctx->alt_keys, &ctx->alt_vals, &ctx->alt2_keys, &ctx->alt2_vals,
                          &ctx->status, &ctx->small, &ctx->aux, &ctx->host_tmp_a,
                          &ctx->host_tmp_b, &ctx->host_tmp_c, &ctx->spg_ws, &ctx->spg_tmp,
                          &ctx->spg_big, &ctx->spg_rank, &ctx->spg_cp, &ctx->spg_plan, &ctx->spg_piece, &ctx->spg_chunk, &ctx->hist_part, &ctx->relin_dev, &ctx->exec_buf[0],
                          &ctx->exec_buf[1], &ctx->exec_buf[2], &ctx->exec_buf[3], &ctx->exec_buf[4],
                          &ctx->exec_buf[5]})
    release(*b);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  prof_clear(ctx);
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->check_event) (void)hipEventDestroy(ctx->check_event);
  delete ctx;
  return LNT_OK;
}

int lnt_context_set_param(lnt_context ctx, const char *name, int64_t value) {
  if (!ctx || !name) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  auto in = [&](int64_t lo, int64_t hi) { return value >= lo && value <= hi; };
  std::string n(name);
  if (n == "sort_items" && (value == 8 || value == 12 || value == 16)) ctx->tuning.items_per_thread = (int)value;
  else if (n == "hist_blocks" && in(1, kMaxHistBlocks)) ctx->tuning.hist_blocks = (int)value;
  else if (n == "merge_items" && (value == 0 || value == 8 || value == 12 || value == 16)) ctx->tuning.merge_items = (int)value;
  else if (n == "merge_block" && (value == 256 || value == 512 || value == 1024)) ctx->tuning.merge_block = (int)value;
  else if (n == "spgemm_small" && in(2, 2048)) ctx->tuning.spgemm_small = (int)value;
  else if (n == "narrow_keys" && in(0, 1)) ctx->tuning.narrow_keys = (int)value;
  else if (n == "row_ranks" && in(0, 1)) ctx->tuning.row_ranks = (int)value;
  else if (n == "spgemm_pieces" && in(0, 1)) ctx->tuning.spgemm_pieces = (int)value;
  else if (n == "spgemm_single_copy" && in(0, 1)) ctx->tuning.spgemm_single_copy = (int)value;
  else if (n == "digit_bits" && in(4, 8)) ctx->tuning.digit_bits = (int)value;
  else if (n == "sort_block" && (value == 256 || value == 512 || value == 1024)) ctx->tuning.sort_block = (int)value;
  else return fail(LNT_ERR_INVALID_ARGUMENT, "unknown parameter or value out of range: " + n);
  return LNT_OK;
}

int lnt_context_get_param(lnt_context ctx, const char *name, int64_t *value) {
  if (!ctx || !name || !value) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  std::string n(name);
  if (n == "sort_items") *value = ctx->tuning.items_per_thread;
  else if (n == "hist_blocks") *value = ctx->tuning.hist_blocks;
  else if (n == "merge_items") *value = ctx->tuning.merge_items;
  else if (n == "merge_block") *value = ctx->tuning.merge_block;
  else if (n == "spgemm_small") *value = ctx->tuning.spgemm_small;
  else if (n == "narrow_keys") *value = ctx->tuning.narrow_keys;
  else if (n == "row_ranks") *value = ctx->tuning.row_ranks;
  else if (n == "spgemm_pieces") *value = ctx->tuning.spgemm_pieces;
  else if (n == "spgemm_single_copy") *value = ctx->tuning.spgemm_single_copy;
  else if (n == "digit_bits") *value = ctx->tuning.digit_bits;
  else if (n == "sort_block") *value = ctx->tuning.sort_block;
  else return fail(LNT_ERR_INVALID_ARGUMENT, "unknown parameter: " + n);
  return LNT_OK;
}

int lnt_context_set_profiling(lnt_context ctx, int on) {
  LNT_CTX(ctx);
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  prof_clear(ctx);
  ctx->profiling = on ? 1 : 0;
  return LNT_OK;
}

int lnt_context_profile(lnt_context ctx, const char *kernel, int64_t *launches,
                        double *total_ms, double *bytes) {
  LNT_CTX(ctx);
  if (!kernel || !launches || !total_ms || !bytes) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  int kind = -1;
  for (int k = 0; k < kProfKinds; ++k)
    if (strcmp(kernel, kProfNames[k]) == 0) kind = k;
  if (kind < 0) return fail(LNT_ERR_INVALID_ARGUMENT, std::string("unknown kernel name ") + kernel);
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  int64_t n = 0;
  double ms = 0, b = 0;
  for (auto &r : ctx->prof) {
    if (r.kind != kind) continue;
    float t = 0;
    if (hipEventElapsedTime(&t, r.start, r.stop) != hipSuccess) {
      (void)hipGetLastError();  // a launch that failed before its stop event
      continue;
    }
    ++n;
    ms += t;
    b += r.bytes;
  }
  *launches = n;
  *total_ms = ms;
  *bytes = b;
  return LNT_OK;
}

int lnt_context_set_stream(lnt_context ctx, void *stream) {
  if (!ctx) return fail(LNT_ERR_INVALID_ARGUMENT, "null context");
  ctx->stream = (hipStream_t)stream;
  return LNT_OK;
}

int lnt_context_synchronize(lnt_context ctx) {
  LNT_CTX(ctx);
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  return LNT_OK;
}

int lnt_context_trim(lnt_context ctx) {
  LNT_CTX(ctx);
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  for (DeviceBuffer *b : {&ctx->alt_keys, &ctx->alt_vals, &ctx->alt2_keys, &ctx->alt2_vals,
                          &ctx->aux, &ctx->host_tmp_a, &ctx->host_tmp_b, &ctx->host_tmp_c,
                          &ctx->spg_ws, &ctx->spg_tmp, &ctx->spg_big, &ctx->spg_rank, &ctx->spg_cp, &ctx->spg_plan, &ctx->spg_piece, &ctx->spg_chunk, &ctx->exec_buf[0], &ctx->exec_buf[1],
                          &ctx->exec_buf[2], &ctx->exec_buf[3], &ctx->exec_buf[4], &ctx->exec_buf[5]})
    release(*b);
  return LNT_OK;
}

int lnt_context_set_validation(lnt_context ctx, int level) {
  if (!ctx) return fail(LNT_ERR_INVALID_ARGUMENT, "null context");
  if (level < 0 || level > 2) return fail(LNT_ERR_INVALID_ARGUMENT, "validation level must be 0, 1 or 2");
  ctx->validate = level;
  return LNT_OK;
}

int lnt_change_linidx(lnt_context ctx, int64_t *keys, int64_t n, int order,
                      const int64_t *dims, const int32_t *cur_seq,
                      const int32_t *new_seq, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  Relin r;
  uint64_t dimensionality;
  LNT_TRY(make_relin(order, dims, cur_seq, new_seq, &r, &dimensionality));
  if (n == 0 || !r.active) return LNT_OK;
  if (!keys) return fail(LNT_ERR_INVALID_ARGUMENT, "null keys");
  if (mem == LNT_MEM_DEVICE) return relin_device_large(ctx, (uint64_t *)keys, (uint64_t)n, r);
  Staging st;
  LNT_TRY(ensure(st.b[0], (size_t)n * 8));
  LNT_HIP(hipMemcpyAsync(st.b[0].ptr, keys, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  LNT_TRY(relin_device_large(ctx, (uint64_t *)st.b[0].ptr, (uint64_t)n, r));
  LNT_HIP(hipMemcpyAsync(keys, st.b[0].ptr, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  return LNT_OK;
}

static int sort_common(lnt_context ctx, int64_t *keys, double *vals, int64_t n,
                       uint64_t bound, const Relin *relin, int mem, uint32_t sorted_low = 0) {
  if (n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (n == 0) return LNT_OK;
  if (!keys || !vals) return fail(LNT_ERR_INVALID_ARGUMENT, "null keys or values");
  // 2^32 entries or more: chunked (sort_device_large).
  auto run = [&](uint64_t *k, double *v) -> int {
    if (n < (int64_t(1) << 32)) return sort_device(ctx, k, v, (uint32_t)n, bound, relin, sorted_low);
    return sort_device_large(ctx, k, v, (uint64_t)n, bound, relin);
  };
  if (mem == LNT_MEM_DEVICE) return run((uint64_t *)keys, vals);
  Staging st;
  LNT_TRY(ensure(st.b[0], (size_t)n * 8));
  LNT_TRY(ensure(st.b[1], (size_t)n * 8));
  LNT_HIP(hipMemcpyAsync(st.b[0].ptr, keys, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  LNT_HIP(hipMemcpyAsync(st.b[1].ptr, vals, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  int rc = run((uint64_t *)st.b[0].ptr, (double *)st.b[1].ptr);
  if (rc != LNT_OK) return rc;
  LNT_HIP(hipMemcpyAsync(keys, st.b[0].ptr, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipMemcpyAsync(vals, st.b[1].ptr, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  return LNT_OK;
}

int lnt_sort_pairs(lnt_context ctx, int64_t *keys, double *vals, int64_t n,
                   uint64_t max_size, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (max_size == 0 || max_size > (uint64_t(1) << 63))
    return fail(LNT_ERR_INVALID_ARGUMENT, "max_size must be in [1, 2^63]");
  return sort_common(ctx, keys, vals, n, max_size, nullptr, mem);
}

int lnt_change_sort_order(lnt_context ctx, int64_t *keys, double *vals,
                          int64_t n, int order, const int64_t *dims,
                          const int32_t *cur_seq, const int32_t *new_seq,
                          int is_sorted, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  Relin r;
  uint64_t dimensionality;
  LNT_TRY(make_relin(order, dims, cur_seq, new_seq, &r, &dimensionality));
  // SparseMIABase.h:590-645: a sorted tensor whose order does not change is
  // left alone; everything else is relinearised and sorted.
  // A sorted tensor already orders the fastest modes of the new sequence
  // when they keep their relative order (the structure RadixShuffle
  // exploits, LibMIARadix.h:733-740): those low bits need no radix pass.
  if (is_sorted && !r.active) return LNT_OK;
  const uint32_t low = is_sorted ? sorted_low_bits(order, dims, cur_seq, new_seq) : 0;
  return sort_common(ctx, keys, vals, n, dimensionality, &r, mem, low);
}

int lnt_compare(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na, const int64_t *b_keys,
                const double *b_vals, int64_t nb, int mode, double precision, int *equal, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!equal) return fail(LNT_ERR_INVALID_ARGUMENT, "null equal");
  if (mode != LNT_CMP_EXACT && mode != LNT_CMP_FUZZY)
    return fail(LNT_ERR_INVALID_ARGUMENT, "mode must be LNT_CMP_EXACT or LNT_CMP_FUZZY");
  if (mode == LNT_CMP_FUZZY && !(precision > 0)) return fail(LNT_ERR_INVALID_ARGUMENT, "precision must be positive");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na >= (int64_t(1) << 32) || nb >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "na, nb must be < 2^32");
  if ((na && (!a_keys || !a_vals)) || (nb && (!b_keys || !b_vals))) return fail(LNT_ERR_INVALID_ARGUMENT, "null array");
  if (mem == LNT_MEM_DEVICE) {
    if (ctx->validate >= 2)
      LNT_TRY(check_sorted(ctx, (const uint64_t *)a_keys, (uint32_t)na, (const uint64_t *)b_keys, (uint32_t)nb));
    return compare_device(ctx, (const uint64_t *)a_keys, a_vals, (uint32_t)na, (const uint64_t *)b_keys, b_vals,
                          (uint32_t)nb, mode, precision, equal);
  }
  Staging st;
  const size_t sa = (size_t)na * 8, sb = (size_t)nb * 8;
  for (int i = 0; i < 4; ++i) LNT_TRY(ensure(st.b[i], (i < 2 ? sa : sb) ? (i < 2 ? sa : sb) : 8));
  if (na) {
    LNT_HIP(hipMemcpyAsync(st.b[0].ptr, a_keys, sa, hipMemcpyHostToDevice, ctx->stream));
    LNT_HIP(hipMemcpyAsync(st.b[1].ptr, a_vals, sa, hipMemcpyHostToDevice, ctx->stream));
  }
  if (nb) {
    LNT_HIP(hipMemcpyAsync(st.b[2].ptr, b_keys, sb, hipMemcpyHostToDevice, ctx->stream));
    LNT_HIP(hipMemcpyAsync(st.b[3].ptr, b_vals, sb, hipMemcpyHostToDevice, ctx->stream));
  }
  if (ctx->validate >= 2)
    LNT_TRY(check_sorted(ctx, (const uint64_t *)st.b[0].ptr, (uint32_t)na, (const uint64_t *)st.b[2].ptr, (uint32_t)nb));
  return compare_device(ctx, (const uint64_t *)st.b[0].ptr, (const double *)st.b[1].ptr, (uint32_t)na,
                        (const uint64_t *)st.b[2].ptr, (const double *)st.b[3].ptr, (uint32_t)nb, mode, precision,
                        equal);
}

int lnt_merge_into(lnt_context ctx, const int64_t *a_keys, const double *a_vals,
                   int64_t na, const int64_t *b_keys, const double *b_vals,
                   int64_t nb, int op, int64_t *c_keys, double *c_vals,
                   int64_t capacity, int64_t *nc, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null nc");
  if (op != LNT_OP_ADD && op != LNT_OP_SUB && op != LNT_OP_HADAMARD)
    return fail(LNT_ERR_INVALID_ARGUMENT, "op must be LNT_OP_ADD, LNT_OP_SUB or LNT_OP_HADAMARD");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na + nb >= (int64_t(1) << 62)) return fail(LNT_ERR_TOO_LARGE, "na + nb must be < 2^62");
  int64_t need = op == LNT_OP_HADAMARD ? std::min(na, nb) : na + nb;
  if (capacity < need)
    return fail(LNT_ERR_INVALID_ARGUMENT, "capacity " + std::to_string(capacity) +
                                              " is below the bound " + std::to_string(need));
  if ((na && (!a_keys || !a_vals)) || (nb && (!b_keys || !b_vals)) || (need && (!c_keys || !c_vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null array");
  // Inputs of 2^32 entries or more go through in merge-path chunks.
  auto run = [&](const uint64_t *ak, const double *av, const uint64_t *bk, const double *bv, uint64_t *ck,
                 double *cv, uint64_t cap, uint64_t *out) -> int {
    if (na + nb < (int64_t(1) << 32)) {
      uint32_t c = 0;
      LNT_TRY(merge_device(ctx, ak, av, (uint32_t)na, bk, bv, (uint32_t)nb, op, ck, cv, cap, &c));
      *out = c;
      return LNT_OK;
    }
    return merge_device_large(ctx, ak, av, (uint64_t)na, bk, bv, (uint64_t)nb, op, ck, cv, cap, out);
  };
  uint64_t count = 0;
  if (mem == LNT_MEM_DEVICE) {
    if (ctx->validate >= 2)
      LNT_TRY(check_sorted(ctx, (const uint64_t *)a_keys, (uint64_t)na, (const uint64_t *)b_keys, (uint64_t)nb));
    LNT_TRY(run((const uint64_t *)a_keys, a_vals, (const uint64_t *)b_keys, b_vals, (uint64_t *)c_keys, c_vals,
                (uint64_t)capacity, &count));
    *nc = (int64_t)count;
    return LNT_OK;
  }
  Staging st;
  size_t sa = (size_t)na * 8, sb = (size_t)nb * 8, sc = (size_t)need * 8;
  LNT_TRY(ensure(st.b[0], sa ? sa : 8));
  LNT_TRY(ensure(st.b[1], sa ? sa : 8));
  LNT_TRY(ensure(st.b[2], sb ? sb : 8));
  LNT_TRY(ensure(st.b[3], sb ? sb : 8));
  LNT_TRY(ensure(st.b[4], sc ? sc : 8));
  LNT_TRY(ensure(st.b[5], sc ? sc : 8));
  if (sa) {
    LNT_HIP(hipMemcpyAsync(st.b[0].ptr, a_keys, sa, hipMemcpyHostToDevice, ctx->stream));
    LNT_HIP(hipMemcpyAsync(st.b[1].ptr, a_vals, sa, hipMemcpyHostToDevice, ctx->stream));
  }
  if (sb) {
    LNT_HIP(hipMemcpyAsync(st.b[2].ptr, b_keys, sb, hipMemcpyHostToDevice, ctx->stream));
    LNT_HIP(hipMemcpyAsync(st.b[3].ptr, b_vals, sb, hipMemcpyHostToDevice, ctx->stream));
  }
  if (ctx->validate >= 2)
    LNT_TRY(check_sorted(ctx, (const uint64_t *)st.b[0].ptr, (uint64_t)na, (const uint64_t *)st.b[2].ptr, (uint64_t)nb));
  LNT_TRY(run((const uint64_t *)st.b[0].ptr, (const double *)st.b[1].ptr, (const uint64_t *)st.b[2].ptr,
              (const double *)st.b[3].ptr, (uint64_t *)st.b[4].ptr, (double *)st.b[5].ptr, (uint64_t)need, &count));
  if (count) {
    LNT_HIP(hipMemcpyAsync(c_keys, st.b[4].ptr, (size_t)count * 8, hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipMemcpyAsync(c_vals, st.b[5].ptr, (size_t)count * 8, hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipStreamSynchronize(ctx->stream));
  }
  *nc = (int64_t)count;
  return LNT_OK;
}

int lnt_merge_into_noalloc(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                           const int64_t *b_keys, const double *b_vals, int64_t nb, int op, int64_t *c_keys,
                           double *c_vals, int64_t capacity, int64_t *nc) {
  LNT_CTX(ctx);
  if (!nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null nc");
  if (op != LNT_OP_ADD && op != LNT_OP_SUB && op != LNT_OP_HADAMARD)
    return fail(LNT_ERR_INVALID_ARGUMENT, "op must be LNT_OP_ADD, LNT_OP_SUB or LNT_OP_HADAMARD");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na + nb >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "na + nb must be < 2^32");
  const int64_t need = op == LNT_OP_HADAMARD ? std::min(na, nb) : na + nb;
  if (capacity < need)
    return fail(LNT_ERR_INVALID_ARGUMENT, "capacity " + std::to_string(capacity) +
                                              " is below the bound " + std::to_string(need));
  if ((na && (!a_keys || !a_vals)) || (nb && (!b_keys || !b_vals)) || (need && (!c_keys || !c_vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null array");
  if (ctx->validate >= 2)
    LNT_TRY(check_sorted(ctx, (const uint64_t *)a_keys, (uint32_t)na, (const uint64_t *)b_keys, (uint32_t)nb));
  uint32_t count = 0;
  LNT_TRY(merge_noalloc_device(ctx, (const uint64_t *)a_keys, a_vals, (uint32_t)na, (const uint64_t *)b_keys, b_vals,
                               (uint32_t)nb, op, (uint64_t *)c_keys, c_vals, (uint64_t)capacity, &count));
  *nc = count;
  return LNT_OK;
}

int lnt_merge(lnt_context ctx, const int64_t *a_keys, const double *a_vals,
              int64_t na, const int64_t *b_keys, const double *b_vals,
              int64_t nb, int op, int64_t **c_keys, double **c_vals,
              int64_t *nc, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!c_keys || !c_vals || !nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  int64_t cap = op == LNT_OP_HADAMARD ? std::min(na, nb) : na + nb;
  size_t bytes = (size_t)(cap > 0 ? cap : 1) * 8;
  void *k = nullptr, *v = nullptr;
  if (mem == LNT_MEM_DEVICE) {
    if (hipMalloc(&k, bytes) != hipSuccess || hipMalloc(&v, bytes) != hipSuccess) {
      (void)hipGetLastError();
      if (k) (void)hipFree(k);
      return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of the merge result failed");
    }
  } else {
    k = malloc(bytes);
    v = malloc(bytes);
    if (!k || !v) {
      free(k);
      free(v);
      return fail(LNT_ERR_OUT_OF_MEMORY, "malloc of the merge result failed");
    }
  }
  int rc = lnt_merge_into(ctx, a_keys, a_vals, na, b_keys, b_vals, nb, op, (int64_t *)k, (double *)v, cap, nc, mem);
  if (rc != LNT_OK) {
    if (mem == LNT_MEM_DEVICE) {
      (void)hipFree(k);
      (void)hipFree(v);
    } else {
      free(k);
      free(v);
    }
    return rc;
  }
  *c_keys = (int64_t *)k;
  *c_vals = (double *)v;
  return LNT_OK;
}

int lnt_sort_merge_into(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                        const int64_t *b_keys, const double *b_vals, int64_t nb, int op, uint64_t bound,
                        int64_t *c_keys, double *c_vals, int64_t capacity, int64_t *nc, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null nc");
  if (op != LNT_OP_ADD && op != LNT_OP_SUB) return fail(LNT_ERR_INVALID_ARGUMENT, "op must be LNT_OP_ADD or LNT_OP_SUB");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na + nb >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "na + nb must be < 2^32");
  if (capacity < na + nb)
    return fail(LNT_ERR_INVALID_ARGUMENT, "capacity " + std::to_string(capacity) + " is below na + nb");
  if ((na && (!a_keys || !a_vals)) || (nb && (!b_keys || !b_vals)) || (na + nb && (!c_keys || !c_vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null array");
  if (bound == 0 || bound > (uint64_t(1) << 63)) return fail(LNT_ERR_INVALID_ARGUMENT, "bound must be in [1, 2^63]");
  *nc = 0;
  uint64_t count = 0;
  if (mem == LNT_MEM_DEVICE) {
    LNT_TRY(sort_merge_device(ctx, (const uint64_t *)a_keys, a_vals, (uint32_t)na, (const uint64_t *)b_keys, b_vals,
                              (uint32_t)nb, op, bound, (uint64_t *)c_keys, c_vals, (uint64_t)capacity, &count));
    *nc = (int64_t)count;
    return LNT_OK;
  }
  Staging st;
  const size_t sa = (size_t)na * 8, sb = (size_t)nb * 8, sc = (size_t)(na + nb) * 8;
  for (int i = 0; i < 2; ++i) LNT_TRY(ensure(st.b[i], sa ? sa : 8));
  for (int i = 2; i < 4; ++i) LNT_TRY(ensure(st.b[i], sb ? sb : 8));
  for (int i = 4; i < 6; ++i) LNT_TRY(ensure(st.b[i], sc ? sc : 8));
  if (sa) {
    LNT_HIP(hipMemcpyAsync(st.b[0].ptr, a_keys, sa, hipMemcpyHostToDevice, ctx->stream));
    LNT_HIP(hipMemcpyAsync(st.b[1].ptr, a_vals, sa, hipMemcpyHostToDevice, ctx->stream));
  }
  if (sb) {
    LNT_HIP(hipMemcpyAsync(st.b[2].ptr, b_keys, sb, hipMemcpyHostToDevice, ctx->stream));
    LNT_HIP(hipMemcpyAsync(st.b[3].ptr, b_vals, sb, hipMemcpyHostToDevice, ctx->stream));
  }
  LNT_TRY(sort_merge_device(ctx, (const uint64_t *)st.b[0].ptr, (const double *)st.b[1].ptr, (uint32_t)na,
                            (const uint64_t *)st.b[2].ptr, (const double *)st.b[3].ptr, (uint32_t)nb, op, bound,
                            (uint64_t *)st.b[4].ptr, (double *)st.b[5].ptr, (uint64_t)(na + nb), &count));
  if (count) {
    LNT_HIP(hipMemcpyAsync(c_keys, st.b[4].ptr, count * 8, hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipMemcpyAsync(c_vals, st.b[5].ptr, count * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  *nc = (int64_t)count;
  return LNT_OK;
}

int lnt_sort_merge(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                   const int64_t *b_keys, const double *b_vals, int64_t nb, int op, uint64_t bound,
                   int64_t **c_keys, double **c_vals, int64_t *nc, int mem) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!c_keys || !c_vals || !nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  const int64_t cap = na + nb;
  const size_t bytes = (size_t)(cap > 0 ? cap : 1) * 8;
  void *k = nullptr, *v = nullptr;
  if (mem == LNT_MEM_DEVICE) {
    if (hipMalloc(&k, bytes) != hipSuccess || hipMalloc(&v, bytes) != hipSuccess) {
      (void)hipGetLastError();
      if (k) (void)hipFree(k);
      return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of the merge result failed");
    }
  } else {
    k = malloc(bytes);
    v = malloc(bytes);
    if (!k || !v) {
      free(k);
      free(v);
      return fail(LNT_ERR_OUT_OF_MEMORY, "malloc of the merge result failed");
    }
  }
  int rc = lnt_sort_merge_into(ctx, a_keys, a_vals, na, b_keys, b_vals, nb, op, bound, (int64_t *)k, (double *)v,
                               cap, nc, mem);
  if (rc != LNT_OK) {
    if (mem == LNT_MEM_DEVICE) {
      (void)hipFree(k);
      (void)hipFree(v);
    } else {
      free(k);
      free(v);
    }
    return rc;
  }
  *c_keys = (int64_t *)k;
  *c_vals = (double *)v;
  return LNT_OK;
}

static int lattice_common(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                          const int64_t a_dims[3], const int64_t *b_keys, const double *b_vals, int64_t nb,
                          const int64_t b_dims[3], int strategy, int64_t *c_keys, double *c_vals,
                          int64_t capacity, int64_t *nc, int64_t **c_keys_alloc, double **c_vals_alloc,
                          int64_t *nprod_only, int mem) {
  LNT_TRY(check_mem(mem));
  if (!a_dims || !b_dims) return fail(LNT_ERR_INVALID_ARGUMENT, "null lattice dimensions");
  if (na < 0 || nb < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "negative length");
  if (na >= (int64_t(1) << 32) || nb >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "operands must hold < 2^32 entries");
  if ((strategy & ~LNT_SPGEMM_CSC) < LNT_SPGEMM_AUTO || (strategy & ~LNT_SPGEMM_CSC) > LNT_SPGEMM_MTT)
    return fail(LNT_ERR_INVALID_ARGUMENT, "unknown strategy");
  if ((na && !a_keys) || (nb && !b_keys) || (!nprod_only && ((na && !a_vals) || (nb && !b_vals))))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null array");
  const uint64_t *ak = (const uint64_t *)a_keys, *bk = (const uint64_t *)b_keys;
  const double *av = a_vals, *bv = b_vals;
  Staging st;
  if (mem == LNT_MEM_HOST) {
    size_t sa = (size_t)na * 8, sb = (size_t)nb * 8;
    LNT_TRY(ensure(st.b[0], sa ? sa : 8));
    LNT_TRY(ensure(st.b[1], sa ? sa : 8));
    LNT_TRY(ensure(st.b[2], sb ? sb : 8));
    LNT_TRY(ensure(st.b[3], sb ? sb : 8));
    if (sa) {
      LNT_HIP(hipMemcpyAsync(st.b[0].ptr, a_keys, sa, hipMemcpyHostToDevice, ctx->stream));
      if (a_vals) LNT_HIP(hipMemcpyAsync(st.b[1].ptr, a_vals, sa, hipMemcpyHostToDevice, ctx->stream));
    }
    if (sb) {
      LNT_HIP(hipMemcpyAsync(st.b[2].ptr, b_keys, sb, hipMemcpyHostToDevice, ctx->stream));
      if (b_vals) LNT_HIP(hipMemcpyAsync(st.b[3].ptr, b_vals, sb, hipMemcpyHostToDevice, ctx->stream));
    }
    ak = (const uint64_t *)st.b[0].ptr;
    av = (const double *)st.b[1].ptr;
    bk = (const uint64_t *)st.b[2].ptr;
    bv = (const double *)st.b[3].ptr;
  }
  if (ctx->validate >= 2) LNT_TRY(check_sorted(ctx, ak, (uint32_t)na, bk, (uint32_t)nb));
  if (nprod_only) {
    uint64_t np = 0;
    LNT_TRY(lattice_nprod_device(ctx, ak, (uint64_t)na, a_dims, bk, (uint64_t)nb, b_dims, &np));
    *nprod_only = (int64_t)np;
    return LNT_OK;
  }
  if ((strategy & ~LNT_SPGEMM_CSC) == LNT_SPGEMM_AUTO) strategy |= LNT_SPGEMM_SORT;
  uint64_t count = 0;
  if (mem == LNT_MEM_DEVICE) {
    uint64_t *kalloc = nullptr;
    double *valloc = nullptr;
    LNT_TRY(lattice_mult_device(ctx, ak, av, (uint64_t)na, a_dims, bk, bv, (uint64_t)nb, b_dims, strategy,
                                (uint64_t *)c_keys, c_vals, (uint64_t)(capacity < 0 ? 0 : capacity), &count,
                                nullptr, c_keys_alloc ? &kalloc : nullptr, c_vals_alloc ? &valloc : nullptr));
    if (c_keys_alloc) {
      if (!kalloc) {  // empty product: hand back a valid 1-element allocation
        if (hipMalloc((void **)&kalloc, 8) != hipSuccess || hipMalloc((void **)&valloc, 8) != hipSuccess)
          return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
      }
      *c_keys_alloc = (int64_t *)kalloc;
      *c_vals_alloc = valloc;
    }
    *nc = (int64_t)count;
    return LNT_OK;
  }
  uint64_t *dk = nullptr;
  double *dv = nullptr;
  LNT_TRY(lattice_mult_device(ctx, ak, av, (uint64_t)na, a_dims, bk, bv, (uint64_t)nb, b_dims, strategy, nullptr,
                              nullptr, 0, &count, nullptr, &dk, &dv));
  int rc = LNT_OK;
  int64_t *hk = c_keys;
  double *hv = c_vals;
  if (c_keys_alloc) {
    hk = (int64_t *)malloc((count ? count : 1) * 8);
    hv = (double *)malloc((count ? count : 1) * 8);
    if (!hk || !hv) {
      free(hk);
      free(hv);
      rc = fail(LNT_ERR_OUT_OF_MEMORY, "malloc of the lattice product failed");
    }
  } else if ((int64_t)count > capacity) {
    rc = fail(LNT_ERR_INVALID_ARGUMENT, "capacity is below nnz(C) = " + std::to_string(count));
  }
  if (rc == LNT_OK && count) {
    if (hipMemcpyAsync(hk, dk, count * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(hv, dv, count * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(LNT_ERR_DEVICE, "copy of the lattice product to the host failed");
  }
  if (dk) (void)hipFree(dk);
  if (dv) (void)hipFree(dv);
  if (rc != LNT_OK) {
    if (c_keys_alloc) {
      free(hk);
      free(hv);
    }
    return rc;
  }
  if (c_keys_alloc) {
    *c_keys_alloc = hk;
    *c_vals_alloc = hv;
  }
  *nc = (int64_t)count;
  return LNT_OK;
}

int lnt_lattice_nprod(lnt_context ctx, const int64_t *a_keys, int64_t na, const int64_t a_dims[3],
                      const int64_t *b_keys, int64_t nb, const int64_t b_dims[3], int64_t *nprod, int mem) {
  LNT_CTX(ctx);
  if (!nprod) return fail(LNT_ERR_INVALID_ARGUMENT, "null nprod");
  return lattice_common(ctx, a_keys, nullptr, na, a_dims, b_keys, nullptr, nb, b_dims, 0, nullptr, nullptr, 0,
                        nullptr, nullptr, nullptr, nprod, mem);
}

int lnt_lattice_mult_into(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                          const int64_t a_dims[3], const int64_t *b_keys, const double *b_vals, int64_t nb,
                          const int64_t b_dims[3], int strategy, int64_t *c_keys, double *c_vals,
                          int64_t capacity, int64_t *nc, int mem) {
  LNT_CTX(ctx);
  if (!nc) return fail(LNT_ERR_INVALID_ARGUMENT, "null nc");
  if (capacity > 0 && (!c_keys || !c_vals)) return fail(LNT_ERR_INVALID_ARGUMENT, "null output array");
  return lattice_common(ctx, a_keys, a_vals, na, a_dims, b_keys, b_vals, nb, b_dims, strategy, c_keys, c_vals,
                        capacity, nc, nullptr, nullptr, nullptr, mem);
}

int lnt_lattice_mult(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                     const int64_t a_dims[3], const int64_t *b_keys, const double *b_vals, int64_t nb,
                     const int64_t b_dims[3], int strategy, int64_t **c_keys, double **c_vals, int64_t *nc,
                     int mem) {
  LNT_CTX(ctx);
  if (!nc || !c_keys || !c_vals) return fail(LNT_ERR_INVALID_ARGUMENT, "null output pointer");
  return lattice_common(ctx, a_keys, a_vals, na, a_dims, b_keys, b_vals, nb, b_dims, strategy, nullptr, nullptr, 0,
                        nc, c_keys, c_vals, nullptr, mem);
}

int lnt_free(lnt_context ctx, void *ptr, int mem) {
  if (!ptr) return LNT_OK;
  if (mem == LNT_MEM_DEVICE) {
    if (ctx) {
      DeviceGuard guard(ctx->device);
      LNT_HIP(hipFree(ptr));
    } else {
      LNT_HIP(hipFree(ptr));
    }
    return LNT_OK;
  }
  if (mem != LNT_MEM_HOST) return fail(LNT_ERR_INVALID_ARGUMENT, "mem must be LNT_MEM_HOST or LNT_MEM_DEVICE");
  free(ptr);
  return LNT_OK;
}

int lnt_copy(lnt_context ctx, void *dst, const void *src, int64_t bytes) {
  LNT_CTX(ctx);
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return fail(LNT_ERR_INVALID_ARGUMENT, "bad copy arguments");
  if (bytes == 0) return LNT_OK;
  LNT_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  return LNT_OK;
}

int lnt_alloc(lnt_context ctx, int64_t bytes, int mem, void **ptr) {
  LNT_CTX(ctx);
  LNT_TRY(check_mem(mem));
  if (!ptr || bytes < 1) return fail(LNT_ERR_INVALID_ARGUMENT, "bytes must be >= 1 and ptr non-null");
  *ptr = nullptr;
  if (mem == LNT_MEM_HOST) {
    *ptr = malloc((size_t)bytes);
    if (!*ptr) return fail(LNT_ERR_OUT_OF_MEMORY, "malloc failed");
    return LNT_OK;
  }
  if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return LNT_OK;
}

}  // extern "C"
