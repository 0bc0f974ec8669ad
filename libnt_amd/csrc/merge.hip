// merge.hip -- merge-path sparse union (+, -) and intersection (Hadamard) of
// two sorted (uint64 key, double value) arrays for gfx950.
//
// Replaces the serial two-pointer merge of the reference
// (multi_index_array/FunctionUtil.h:373-419 outside_merge_sparse_storage_containers)
// and the pure Hadamard case of SparseMIABase::perform_noLatticeMult
// (SparseMIABase.h:1717-1826).  Value rules, exactly as the reference:
//   union:        equal keys -> op(a, b); A only -> a; B only -> op(0, b)
//   intersection: equal keys -> a * b
// Each value is a single IEEE operation, so results are bit-identical.
//
// One kernel, one pass over the inputs.  The merge path of A and B (ties: A
// first) is cut into tiles of 256*ITEMS steps; a tile finds its end points by
// binary search on its two diagonals, loads its slices into LDS, and each
// thread merges ITEMS consecutive steps.  Because ties put A first, a key
// present in both inputs appears as A[i] immediately followed by B[j]: the A
// step emits the combined value (peeking at the next B key, which may belong to
// the next tile) and the B step is dropped (peeking at the previous A key,
// which may belong to the previous tile).  Output counts are block-scanned and
// the tile's output offset comes from a decoupled look-back; outputs are staged
// in LDS and written contiguously.
#include <string.h>

#include <algorithm>

#include "lnt_internal.h"

namespace lnt {

namespace {

constexpr int kBlock = 256;

// Number of A elements among the first `diag` steps of the merge path.
__device__ inline uint32_t path_search(const uint64_t *a, uint32_t na,
                                       const uint64_t *b, uint32_t nb,
                                       uint32_t diag) {
  uint32_t lo = diag > nb ? diag - nb : 0;
  uint32_t hi = diag < na ? diag : na;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= b[diag - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Merge-path split of every tile boundary, all searched in parallel; also
// resets the tile counter of the merge kernel that follows.
__global__ void partition_kernel(const uint64_t *__restrict__ a, uint32_t na,
                                 const uint64_t *__restrict__ b, uint32_t nb,
                                 uint32_t tile, uint32_t num_tiles,
                                 uint32_t *__restrict__ splits,
                                 uint32_t *__restrict__ tile_counter) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *tile_counter = 0;
  if (t > num_tiles) return;
  uint64_t total = (uint64_t)na + nb;
  uint32_t diag = (uint32_t)min<uint64_t>((uint64_t)t * tile, total);
  splits[t] = path_search(a, na, b, nb, diag);
}

template <int ITEMS>
__global__ __launch_bounds__(kBlock) void merge_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, uint32_t na,
    const uint64_t *__restrict__ bk, const double *__restrict__ bv, uint32_t nb,
    int op, uint64_t *__restrict__ ck, double *__restrict__ cv,
    const uint32_t *__restrict__ splits, uint64_t *__restrict__ status,
    uint32_t epoch, uint32_t *__restrict__ tile_counter,
    uint32_t *__restrict__ total_out, uint32_t num_tiles, uint64_t capacity) {
  constexpr int kTile = kBlock * ITEMS;
  __shared__ uint64_t s_key[kTile + 1];
  __shared__ double s_val[kTile + 1];
  __shared__ uint32_t s_tile, s_i0, s_i1;
  __shared__ uint64_t s_prev_a;
  __shared__ uint32_t s_has_prev;
  __shared__ uint32_t s_scan[4];
  __shared__ uint32_t s_excl;

  const uint32_t tid = threadIdx.x;
  const uint64_t total = (uint64_t)na + nb;
  if (tid == 0)
    s_tile = __hip_atomic_fetch_add(tile_counter, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t d0 = (uint32_t)min<uint64_t>((uint64_t)tile * kTile, total);
  const uint32_t d1 = (uint32_t)min<uint64_t>((uint64_t)d0 + kTile, total);
  if (tid == 0) s_i0 = splits[tile];
  if (tid == 1) s_i1 = splits[tile + 1];
  __syncthreads();
  const uint32_t i0 = s_i0, i1 = s_i1;
  const uint32_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t la = i1 - i0, lb = j1 - j0;
  // A slice, B slice, then the B element after the slice (if any).  All of a
  // thread's loads are issued before any is used, so a tile waits for one
  // memory round trip, not ITEMS of them.
  {
    uint64_t lk[ITEMS];
    double lv[ITEMS];
    const uint32_t last = la + lb - 1;  // a tile holds at least one step
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      uint32_t x = (uint32_t)u * kBlock + tid;
      x = x < la + lb ? x : last;  // clamped, unconditional loads
      const bool from_a = x < la;
      const uint64_t *kp = from_a ? ak + i0 + x : bk + j0 + (x - la);
      const double *vp = from_a ? av + i0 + x : bv + j0 + (x - la);
      lk[u] = *kp;
      lv[u] = *vp;
    }
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      uint32_t x = (uint32_t)u * kBlock + tid;
      if (x < la + lb) {
        s_key[x] = lk[u];
        s_val[x] = lv[u];
      }
    }
  }
  const bool has_next_b = j1 < nb;
  if (tid == 0) {
    if (has_next_b) {
      s_key[la + lb] = bk[j1];
      s_val[la + lb] = bv[j1];
    }
    s_has_prev = i0 > 0;
    s_prev_a = i0 > 0 ? ak[i0 - 1] : 0;
  }
  __syncthreads();
  const uint64_t *sa = s_key, *sb = s_key + la;
  const double *sav = s_val, *sbv = s_val + la;
  const uint32_t nlb = lb + (has_next_b ? 1 : 0);  // B keys visible to peeks

  // Per-thread sub-path.
  const uint32_t t0 = min<uint32_t>(tid * ITEMS, la + lb);
  const uint32_t t1 = min<uint32_t>(t0 + ITEMS, la + lb);
  uint32_t ia, jb;
  {
    uint32_t lo = t0 > lb ? t0 - lb : 0, hi = t0 < la ? t0 : la;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (sa[mid] <= sb[t0 - 1 - mid]) lo = mid + 1;
      else hi = mid;
    }
    ia = lo;
    jb = t0 - lo;
  }
  // Step s of this thread fills slot s; emit marks the slots that produce
  // an output (register arrays are only indexed by unrolled constants).
  uint64_t ok[ITEMS];
  double ov[ITEMS];
  uint32_t emit = 0;
#pragma unroll
  for (int s = 0; s < ITEMS; ++s) {
    ok[s] = 0;
    ov[s] = 0.0;
    if (t0 + s < t1) {
      bool take_a = ia < la && (jb >= lb || sa[ia] <= sb[jb]);
      if (take_a) {
        uint64_t key = sa[ia];
        double a = sav[ia];
        bool match = jb < nlb && sb[jb] == key;
        double b = match ? sbv[jb] : 0.0;
        ok[s] = key;
        if (op == LNT_OP_HADAMARD) {
          ov[s] = a * b;
          emit |= (match ? 1u : 0u) << s;
        } else {
          ov[s] = match ? (op == LNT_OP_SUB ? a - b : a + b) : a;
          emit |= 1u << s;
        }
        ++ia;
      } else {
        uint64_t key = sb[jb];
        bool absorbed = ia > 0 ? sa[ia - 1] == key : (s_has_prev && s_prev_a == key);
        ok[s] = key;
        ov[s] = op == LNT_OP_SUB ? 0.0 - sbv[jb] : 0.0 + sbv[jb];
        emit |= ((!absorbed && op != LNT_OP_HADAMARD) ? 1u : 0u) << s;
        ++jb;
      }
    }
  }
  const uint32_t cnt = (uint32_t)__popc(emit);
  uint32_t tile_count;
  uint32_t my_off = block256_exclusive_sum(cnt, s_scan, &tile_count);

  if (tid < kWave) {  // wave 0 publishes and looks back
    uint32_t excl = 0;
    uint64_t *my_status = status + tile;
    if (tile == 0) {
      if (tid == 0) status_store(my_status, status_word(kFlagInclusive, epoch, tile_count));
    } else {
      if (tid == 0) status_store(my_status, status_word(kFlagAggregate, epoch, tile_count));
      excl = wave_lookback(status, tile, epoch);
      if (tid == 0) status_store(my_status, status_word(kFlagInclusive, epoch, excl + tile_count));
    }
    if (tid == 0) {
      if (tile == num_tiles - 1) *total_out = excl + tile_count;
      s_excl = excl;
    }
  }
  __syncthreads();  // everyone is done reading the input slices
#pragma unroll
  for (int s = 0; s < ITEMS; ++s)
    if ((emit >> s) & 1u) {
      uint32_t idx = my_off + (uint32_t)__popc(emit & ((1u << s) - 1u));
      s_key[idx] = ok[s];
      s_val[idx] = ov[s];
    }
  __syncthreads();
  const uint64_t base = s_excl;
  // Sorted unique inputs never exceed the capacity; malformed ones (duplicate
  // keys) could, so writes are bounded and the host reports the overflow.
  for (uint32_t x = tid; x < tile_count && base + x < capacity; x += kBlock) {
    ck[base + x] = s_key[x];
    cv[base + x] = s_val[x];
  }
}

template <int ITEMS>
int launch_merge(lnt_context ctx, const uint64_t *ak, const double *av,
                 uint32_t na, const uint64_t *bk, const double *bv, uint32_t nb,
                 int op, uint64_t *ck, double *cv, uint64_t capacity,
                 uint32_t *counter, uint32_t *total, int *slot) {
  constexpr uint64_t tile = kBlock * ITEMS;
  uint64_t steps = (uint64_t)na + nb;
  uint32_t tiles = (uint32_t)((steps + tile - 1) / tile);
  LNT_TRY(ensure_status(ctx, tiles));
  LNT_TRY(ensure(ctx->aux, ((size_t)tiles + 1) * sizeof(uint32_t)));
  uint32_t *splits = (uint32_t *)ctx->aux.ptr;
  uint32_t epoch;
  LNT_TRY(next_epoch(ctx, &epoch));
  int pslot = prof_begin(ctx, kProfPartition, 0.0);
  partition_kernel<<<(tiles + 1 + 255) / 256, 256, 0, ctx->stream>>>(
      ak, na, bk, nb, (uint32_t)tile, tiles, splits, counter);
  prof_end(ctx, pslot);
  LNT_HIP(hipGetLastError());
  *slot = prof_begin(ctx, kProfMerge, 16.0 * ((double)na + nb));
  merge_kernel<ITEMS><<<tiles, kBlock, 0, ctx->stream>>>(
      ak, av, na, bk, bv, nb, op, ck, cv, splits, (uint64_t *)ctx->status.ptr,
      epoch, counter, total, tiles, capacity);
  prof_end(ctx, *slot);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

// ------------------------------------------------------------ sortMerge
// SparseMIA::sortMerge (SparseMIA.h:833-887): A then B concatenated (B
// negated, -1*b, for subtraction), sorted, then collect_duplicates
// (SparseMIA.h:628-665).
__global__ void concat_kernel(const uint64_t *__restrict__ ak, const double *__restrict__ av, uint32_t na,
                              const uint64_t *__restrict__ bk, const double *__restrict__ bv, uint32_t nb,
                              int negate_b, uint64_t *__restrict__ k, double *__restrict__ v) {
  const uint64_t n = (uint64_t)na + nb;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < na) {
      k[i] = ak[i];
      v[i] = av[i];
    } else {
      k[i] = bk[i - na];
      v[i] = negate_b ? -1.0 * bv[i - na] : bv[i - na];
    }
  }
}

// collect_duplicates on sorted keys: one thread per run of equal keys (run
// heads).  The run's first entry is kept only if its value is non-zero (the
// reference tests `if(this->data_at(first))` before appending a new key) --
// except the first entry of the whole array, which is always kept; a zero
// head is skipped and the next entry of the same key is tried as a new key;
// entries after the kept one are added in sorted order.  keep[p] / v[p]
// receive the run's outcome at its head; the sum overwrites v[p] in place
// (runs are disjoint).
__global__ void collect_kernel(const uint64_t *__restrict__ k, double *__restrict__ v, uint32_t n,
                               uint32_t *__restrict__ keep) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = k[p];
    if (p > 0 && k[p - 1] == key) {
      keep[p] = 0;
      continue;
    }
    uint64_t q = p;
    if (p > 0)
      while (q < n && k[q] == key && v[q] == 0.0) ++q;
    if (q == n || k[q] != key) {  // every entry of the key is zero: dropped
      keep[p] = 0;
      continue;
    }
    double sum = v[q];
    for (uint64_t r = q + 1; r < n && k[r] == key; ++r) sum = sum + v[r];
    v[p] = sum;
    keep[p] = 1;
  }
}

__global__ void compact_kernel(const uint64_t *__restrict__ k, const double *__restrict__ v,
                               const uint32_t *__restrict__ keep, const uint64_t *__restrict__ pos, uint32_t n,
                               uint64_t *__restrict__ ck, double *__restrict__ cv, uint64_t capacity) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
    if (!keep[p]) continue;
    const uint64_t d = pos[p];
    if (d < capacity) {
      ck[d] = k[p];
      cv[d] = v[p];
    }
  }
}

}  // namespace

int merge_device(lnt_context ctx, const uint64_t *ak, const double *av,
                 uint32_t na, const uint64_t *bk, const double *bv, uint32_t nb,
                 int op, uint64_t *ck, double *cv, uint64_t capacity,
                 uint32_t *nc_host) {
  if ((uint64_t)na + nb == 0) {
    *nc_host = 0;
    return LNT_OK;
  }
  LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
  uint32_t *small = (uint32_t *)ctx->small.ptr;
  uint32_t *counter = small + kSmallCounters + 15;
  uint32_t *total = small + kSmallFlags + 1;
  int rc, slot = -1;
  if (ctx->tuning.merge_items == 16)
    rc = launch_merge<16>(ctx, ak, av, na, bk, bv, nb, op, ck, cv, capacity, counter, total, &slot);
  else
    rc = launch_merge<8>(ctx, ak, av, na, bk, bv, nb, op, ck, cv, capacity, counter, total, &slot);
  if (rc != LNT_OK) return rc;
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 1, total, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  *nc_host = ctx->pinned[1];
  prof_set_bytes(ctx, slot, 16.0 * ((double)na + nb + *nc_host));
  if (*nc_host > capacity)
    return fail(LNT_ERR_NOT_SORTED, "merge inputs are not sorted with unique keys (output overflow)");
  return LNT_OK;
}

template <typename T>
int exclusive_scan(lnt_context ctx, const T *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                   DeviceBuffer &tmp, size_t tmp_offset);
size_t scan_workspace_words(uint64_t n);

// SparseMIA::sortMerge of two operands in one linearisation, keys < bound.
// Writes at most capacity (>= na + nb) entries; *nc_host = nnz(C).
int sort_merge_device(lnt_context ctx, const uint64_t *ak, const double *av, uint32_t na, const uint64_t *bk,
                      const double *bv, uint32_t nb, int op, uint64_t bound, uint64_t *ck, double *cv,
                      uint64_t capacity, uint64_t *nc_host) {
  const uint64_t n = (uint64_t)na + nb;
  *nc_host = 0;
  if (n == 0) return LNT_OK;
  if (n >= (uint64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, "na + nb must be < 2^32");
  // Work arrays: the concatenation (keys, values), keep flags, positions.
  const size_t words = scan_workspace_words(n) + 64;
  LNT_TRY(ensure(ctx->exec_buf[0], n * 8));
  LNT_TRY(ensure(ctx->exec_buf[1], n * 8));
  LNT_TRY(ensure(ctx->exec_buf[2], n * 4 + 256));
  LNT_TRY(ensure(ctx->exec_buf[3], (n + 1) * 8 + 256));
  LNT_TRY(ensure(ctx->aux, words * 8));
  uint64_t *k = (uint64_t *)ctx->exec_buf[0].ptr;
  double *v = (double *)ctx->exec_buf[1].ptr;
  uint32_t *keep = (uint32_t *)ctx->exec_buf[2].ptr;
  uint64_t *pos = (uint64_t *)ctx->exec_buf[3].ptr;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
  concat_kernel<<<grid, 256, 0, ctx->stream>>>(ak, av, na, bk, bv, nb, op == LNT_OP_SUB, k, v);
  LNT_HIP(hipGetLastError());
  LNT_TRY(sort_device(ctx, k, v, (uint32_t)n, bound, nullptr));
  collect_kernel<<<grid, 256, 0, ctx->stream>>>(k, v, (uint32_t)n, keep);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, keep, pos, n, pos + n, ctx->aux, 0));
  compact_kernel<<<grid, 256, 0, ctx->stream>>>(k, v, keep, pos, (uint32_t)n, ck, cv, capacity);
  LNT_HIP(hipGetLastError());
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 8, pos + n, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  uint64_t nc;
  memcpy(&nc, ctx->pinned + 8, sizeof(nc));
  *nc_host = nc;
  return LNT_OK;
}

}  // namespace lnt
