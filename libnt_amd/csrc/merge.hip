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
// first) is cut into tiles of BT*ITEMS steps; a tile finds its end points in
// the splits precomputed by partition_kernel and loads only the KEYS of its two
// slices into LDS, as 32-bit words relative to the tile's smallest key (64-bit
// keys split into two word arrays in the rare tile whose keys span 2^32).  Each thread merges ITEMS consecutive steps with the two
// current keys held in registers (one LDS load per step) and records, per step,
// a descriptor: which A entry and/or which B entry produces the output.
// Because ties put A first, a key present in both inputs appears as A[i]
// immediately followed by B[j]: the A step emits the combined entry (peeking
// at the next B key, which may belong to the next tile) and the B step is
// dropped (comparing with the last A key, which may belong to the previous
// tile).  Output counts are block-scanned, the tile's output offset comes from
// a decoupled look-back, and the emitted descriptors are compacted in LDS.
// The write phase runs in output order: thread x of the block handles outputs
// x, x+BT, ...; it reads the key from LDS and gathers the value(s) from HBM --
// neighbouring lanes read neighbouring entries of A and B, so the value reads
// coalesce and every value is read once (a Hadamard tile reads only the values
// of matched keys).
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

#include "lnt_internal.h"

namespace lnt {

namespace {


// Number of A elements among the first `diag` steps of the merge path (ties:
// A first), found by G lanes of a wave in rounds of G probes: the predicate
// a[x] <= b[diag-1-x] holds for x below the answer and fails from it on, so
// one ballot over G evenly spaced x narrows the range (G+1)-fold -- with
// G = 4, eleven rounds of parallel loads for 2^26 steps where a binary
// search needs twenty-six dependent ones, at less than twice its cache
// lines.  Every lane of the group returns the answer; the loop runs until
// every group of the wave is done.
template <uint32_t G>
__device__ inline uint32_t group_path_search(const uint64_t *__restrict__ a, uint32_t na,
                                             const uint64_t *__restrict__ b, uint32_t nb, uint32_t diag) {
  static_assert(G >= 2 && G <= kWave && (kWave % G) == 0, "lanes per search");
  const uint32_t lane = threadIdx.x & (kWave - 1), gl = lane % G, shift = lane - gl;
  uint32_t lo = diag > nb ? diag - nb : 0;
  uint32_t hi = diag < na ? diag : na;
  while (__ballot(hi > lo)) {
    const uint32_t w = hi - lo;
    // Probe x: lo + gl when the range fits the group, else strictly
    // increasing points inside [lo, hi).  Every x is in [lo, hi), so
    // x < na and diag - 1 - x < nb.
    const uint32_t x = w <= G ? lo + gl : lo + (uint32_t)(((uint64_t)(gl + 1) * w) / (G + 1));
    const bool live = w > G ? true : gl < w;
    const bool p = live && a[x] <= b[diag - 1 - x];
    const uint64_t mask = G == kWave ? ~0ull : ((1ull << G) - 1);
    const uint32_t trues = (uint32_t)__popcll((__ballot(p) >> shift) & mask);
    // Lanes [0, trues) of the group hold; the answer lies after the last
    // holding probe and at or before the first failing one.
    const uint32_t x_last = __shfl(x, shift + (trues ? trues - 1 : 0), kWave);
    const uint32_t x_fail = __shfl(x, shift + (trues < G ? trues : G - 1), kWave);
    if (w == 0) continue;
    if (w <= G) {
      lo += trues;
      hi = lo;
    } else {
      if (trues) lo = x_last + 1;
      if (trues < G) hi = x_fail;
    }
  }
  return lo;
}

constexpr uint32_t kSplitLanes = 4;  // lanes per merge-path search

// Merge-path split of every tile boundary, kSplitLanes lanes per boundary,
// all searched in parallel; also resets the tile counter of the merge kernel
// that follows.
__global__ __launch_bounds__(256) void partition_kernel(const uint64_t *__restrict__ a, uint32_t na,
                                                        const uint64_t *__restrict__ b, uint32_t nb, uint32_t tile,
                                                        uint32_t num_tiles, uint32_t *__restrict__ splits,
                                                        uint32_t *__restrict__ tile_counter) {
  const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) / kSplitLanes;
  if (blockIdx.x == 0 && threadIdx.x == 0) *tile_counter = 0;
  // Boundaries past the last one search the final diagonal and store nothing,
  // so that every lane of a wave takes part in its ballots.
  const uint64_t total = (uint64_t)na + nb;
  const uint32_t diag = (uint32_t)min<uint64_t>((uint64_t)min(t, num_tiles) * tile, total);
  const uint32_t split = group_path_search<kSplitLanes>(a, na, b, nb, diag);
  if (t <= num_tiles && (threadIdx.x % kSplitLanes) == 0) splits[t] = split;
}

// Step descriptor: A entry index [14:0], B entry index [29:15], flags.
constexpr uint32_t kDescA = 1u << 31;
constexpr uint32_t kDescB = 1u << 30;
constexpr uint32_t kDescIdx = 0x7FFFu;

// Keys of one tile's A slice then B slice into registers: item u of thread
// tid is merge-path step u*BT + tid of the tile.  All loads are issued before
// any is used; a tile holds at least one step, so clamped loads stay inside it.
template <int BT, int ITEMS>
__device__ inline void load_tile_keys(const uint64_t *__restrict__ ak, const uint64_t *__restrict__ bk,
                                      uint32_t i0, uint32_t la, uint32_t j0, uint32_t lb,
                                      uint64_t (&lk)[ITEMS]) {
  const uint32_t last = la + lb - 1;
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    uint32_t x = (uint32_t)u * BT + threadIdx.x;
    x = x < la + lb ? x : last;
    lk[u] = x < la ? ak[i0 + x] : bk[j0 + (x - la)];
  }
}

// A tile keeps its keys in LDS as 32-bit words.  When the keys of both slices
// span less than kNoMatch -- the usual case: a tile covers BT*ITEMS
// consecutive merged keys -- each word is key - base, base being the tile's
// smallest key, and kNoMatch stands for a key outside the tile that equals
// none of its keys.  Otherwise (a wide tile) the low words go to s_lo and the
// high words to s_hi, the descriptor array, which is free until the merge and
// the count scan are done.  Either way the choice is the tile's own: no host
// round trip decides it.
constexpr uint32_t kNoMatch = 0xFFFFFFFFu;

// Sequential merge of the steps [t0, t1) of the tile, one thread.  Slot s of
// desc gets step t0 + s's descriptor; the returned mask marks the slots that
// produce an output.  prev is the A key before the tile (as a tile word in a
// narrow tile), used only if has_prev.
template <int ITEMS, bool WIDE>
__device__ inline uint32_t merge_steps(const uint32_t *__restrict__ s_lo, const uint32_t *__restrict__ s_hi,
                                       uint32_t la, uint32_t lb, uint32_t nlb, uint32_t t0, uint32_t t1,
                                       bool has_prev, uint64_t prev, bool hadamard, uint32_t (&desc)[ITEMS]) {
  using K = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
  auto key = [&](uint32_t x) -> K {
    if constexpr (WIDE) return ((uint64_t)s_hi[x] << 32) | s_lo[x];
    else return s_lo[x];
  };
  // Per-thread sub-path: the B slice starts at word la.
  uint32_t ia, jb;
  {
    uint32_t lo = t0 > lb ? t0 - lb : 0, hi = t0 < la ? t0 : la;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (key(mid) <= key(la + t0 - 1 - mid)) lo = mid + 1;
      else hi = mid;
    }
    ia = lo;
    jb = t0 - lo;
  }
  K ka = ia < la ? key(ia) : 0;
  K kb = jb < nlb ? key(la + jb) : 0;
  bool has_last = ia > 0 || has_prev;  // an A key precedes this sub-path
  K last_a = ia > 0 ? key(ia - 1) : (K)prev;
  uint32_t emit = 0;
#pragma unroll
  for (int s = 0; s < ITEMS; ++s) {
    desc[s] = 0;
    if (t0 + s < t1) {
      const bool take_a = ia < la && (jb >= lb || ka <= kb);
      if (take_a) {
        const bool match = jb < nlb && kb == ka;
        desc[s] = kDescA | ia | (match ? (kDescB | (jb << 15)) : 0u);
        emit |= (uint32_t)(match || !hadamard) << s;
        last_a = ka;
        has_last = true;
        ++ia;
      } else {
        const bool absorbed = has_last && last_a == kb;
        desc[s] = kDescB | (jb << 15);
        emit |= (uint32_t)(!absorbed && !hadamard) << s;
        ++jb;
      }
      // One key read per step: the next key of the side just consumed.  The
      // index never exceeds la + lb (a B step happens only while jb < lb), so
      // it stays inside the arrays; a key read past its slice is never used
      // (the ia < la and jb < nlb tests guard every use).
      const K nk = key(take_a ? ia : la + jb);
      if (take_a) ka = nk;
      else kb = nk;
    }
  }
  return emit;
}

// force_wide: every tile takes the 64-bit path (a test and tuning switch).
template <int BT, int ITEMS>
__global__ __launch_bounds__(BT) void merge_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, uint32_t na,
    const uint64_t *__restrict__ bk, const double *__restrict__ bv, uint32_t nb,
    int op, uint64_t *__restrict__ ck, double *__restrict__ cv,
    const uint32_t *__restrict__ splits, uint64_t *__restrict__ status,
    uint32_t epoch, uint32_t *__restrict__ tile_counter,
    uint32_t *__restrict__ total_out, uint32_t num_tiles, uint64_t capacity, int force_wide) {
  constexpr int kTile = BT * ITEMS;
  static_assert(kTile <= (int)kDescIdx, "descriptor indices are 15 bits");
  static_assert(ITEMS <= 32, "emit mask is 32 bits");
  __shared__ uint32_t s_key[kTile + 1];   // key words (low words of a wide tile)
  __shared__ uint32_t s_desc[kTile + 1];  // high words of a wide tile, then descriptors
  __shared__ uint32_t s_tile, s_i0, s_i1;
  __shared__ uint32_t s_scan[BT / kWave];
  __shared__ uint32_t s_excl;

  const uint32_t tid = threadIdx.x;
  const uint64_t total = (uint64_t)na + nb;
  // Tiles are numbered in the order workgroups start (a global counter), so a
  // tile's look-back only waits on tiles that are already running.
  if (tid == 0) {
    const uint32_t t = __hip_atomic_fetch_add(tile_counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_tile = t;
    s_i0 = splits[t];
    s_i1 = splits[t + 1];
  }
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t i0 = s_i0, i1 = s_i1;
  const uint32_t d0 = (uint32_t)min<uint64_t>((uint64_t)tile * kTile, total);
  const uint32_t d1 = (uint32_t)min<uint64_t>((uint64_t)d0 + kTile, total);
  const uint32_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t la = i1 - i0, lb = j1 - j0;  // la + lb >= 1: a tile holds a step
  const bool has_next_b = j1 < nb;
  // The tile's smallest and largest keys are the first and last ones of its
  // slices; the B key after the slice and the A key before it (read by every
  // thread, alongside the slice loads: the same cache lines).  The next B key
  // is at least the base (it follows every step of the tile) and the previous
  // A key at most the base (it precedes every step).
  const uint64_t lo_key = la ? (lb ? min(ak[i0], bk[j0]) : ak[i0]) : bk[j0];
  const uint64_t hi_key = la ? (lb ? max(ak[i1 - 1], bk[j1 - 1]) : ak[i1 - 1]) : bk[j1 - 1];
  const uint64_t next_b = has_next_b ? bk[j1] : 0;
  const uint64_t prev_a = i0 > 0 ? ak[i0 - 1] : 0;
  uint64_t lk[ITEMS];
  load_tile_keys<BT, ITEMS>(ak, bk, i0, la, j0, lb, lk);
  const bool wide = force_wide || hi_key - lo_key >= kNoMatch;
  if (wide) {
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      const uint32_t x = (uint32_t)u * BT + tid;
      if (x < la + lb) {
        s_key[x] = (uint32_t)lk[u];
        s_desc[x] = (uint32_t)(lk[u] >> 32);
      }
    }
    if (tid == 0 && has_next_b) {
      s_key[la + lb] = (uint32_t)next_b;
      s_desc[la + lb] = (uint32_t)(next_b >> 32);
    }
  } else {
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      const uint32_t x = (uint32_t)u * BT + tid;
      if (x < la + lb) s_key[x] = (uint32_t)(lk[u] - lo_key);
    }
    if (tid == 0 && has_next_b) s_key[la + lb] = (uint32_t)min<uint64_t>(next_b - lo_key, kNoMatch);
  }
  __syncthreads();
  const uint32_t nlb = lb + (has_next_b ? 1 : 0);  // B keys visible to peeks
  const bool hadamard = op == LNT_OP_HADAMARD;
  const uint32_t t0 = min<uint32_t>(tid * ITEMS, la + lb);
  const uint32_t t1 = min<uint32_t>(t0 + ITEMS, la + lb);
  uint32_t desc[ITEMS];
  const uint32_t emit =
      wide ? merge_steps<ITEMS, true>(s_key, s_desc, la, lb, nlb, t0, t1, i0 > 0, prev_a, hadamard, desc)
           : merge_steps<ITEMS, false>(s_key, s_desc, la, lb, nlb, t0, t1, i0 > 0,
                                       prev_a == lo_key ? 0u : kNoMatch, hadamard, desc);
  const uint32_t cnt = (uint32_t)__popc(emit);
  uint32_t tile_count;
  // The scan's barriers also end every thread's reads of s_desc as high words.
  const uint32_t my_off = block_exclusive_sum_bt<BT>(cnt, s_scan, &tile_count);

  // The tile's count is published at once; the look-back for its output
  // offset runs while the value gathers are in flight (only the stores need
  // the offset).
  uint64_t *my_status = status + tile;
  if (tid == 0)
    status_store(my_status, status_word(tile == 0 ? kFlagInclusive : kFlagAggregate, epoch, tile_count));
#pragma unroll
  for (int s = 0; s < ITEMS; ++s)
    if ((emit >> s) & 1u) s_desc[my_off + (uint32_t)__popc(emit & ((1u << s) - 1u))] = desc[s];
  __syncthreads();
  // Write phase in output order: descriptors, then every gather of the
  // thread, then (after the offset is known) the stores.
  const double *tav = av + i0, *tbv = bv + j0;
  uint32_t dd[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    const uint32_t x = (uint32_t)u * BT + tid;
    dd[u] = x < tile_count ? s_desc[x] : 0u;
  }
  double va[ITEMS], vb[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    va[u] = (dd[u] & kDescA) ? tav[dd[u] & kDescIdx] : 0.0;
    vb[u] = (dd[u] & kDescB) ? tbv[(dd[u] >> 15) & kDescIdx] : 0.0;
  }
  if (tid < kWave) {  // wave 0 looks back
    const uint32_t excl = tile == 0 ? 0u : wave_lookback(status, tile, epoch);
    if (tid == 0) {
      if (tile > 0) status_store(my_status, status_word(kFlagInclusive, epoch, excl + tile_count));
      if (tile == num_tiles - 1) *total_out = excl + tile_count;
      s_excl = excl;
    }
  }
  __syncthreads();
  const uint64_t base = s_excl;
  // Sorted unique inputs never exceed the capacity; malformed ones (duplicate
  // keys) could, so writes are bounded and the host reports the overflow.
  const uint32_t lim = (uint32_t)min<uint64_t>(tile_count, capacity > base ? capacity - base : 0);
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    const uint32_t x = (uint32_t)u * BT + tid;
    if (x < lim) {
      const uint32_t d = dd[u];
      const bool has_a = (d & kDescA) != 0, has_b = (d & kDescB) != 0;
      const double a = va[u], b = vb[u];
      double v;
      if (has_a && has_b)
        v = hadamard ? a * b : (op == LNT_OP_SUB ? a - b : a + b);
      else if (has_a)
        v = a;
      else
        v = op == LNT_OP_SUB ? 0.0 - b : 0.0 + b;
      if (!wide) ck[base + x] = lo_key + s_key[has_a ? (d & kDescIdx) : la + ((d >> 15) & kDescIdx)];
      cv[base + x] = v;
    }
  }
  if (wide) {
    // A wide tile's high words are gone from LDS: its keys are read again
    // from the slices (lines the tile has just read), after the value stores
    // so that the loads take the registers the values held.
    const uint64_t *tak = ak + i0, *tbk = bk + j0;
    uint64_t kk[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      const uint32_t x = (uint32_t)u * BT + tid;
      const uint32_t d = dd[u];
      kk[u] = x < lim ? ((d & kDescA) ? tak[d & kDescIdx] : tbk[(d >> 15) & kDescIdx]) : 0;
    }
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      const uint32_t x = (uint32_t)u * BT + tid;
      if (x < lim) ck[base + x] = kk[u];
    }
  }
}

// Device scratch of one merge: look-back words (one per tile), merge-path
// splits (tiles + 1), the tile counter and the output total.
struct MergeScratch {
  uint64_t *status;
  uint32_t *splits, *counter, *total;
  uint32_t epoch;
};

template <int BT, int ITEMS>
int launch_merge(lnt_context ctx, const uint64_t *ak, const double *av,
                 uint32_t na, const uint64_t *bk, const double *bv, uint32_t nb,
                 int op, uint64_t *ck, double *cv, uint64_t capacity,
                 const MergeScratch &m, int *slot) {
  constexpr uint64_t tile = BT * ITEMS;
  uint64_t steps = (uint64_t)na + nb;
  uint32_t tiles = (uint32_t)((steps + tile - 1) / tile);
  int pslot = prof_begin(ctx, kProfPartition, 0.0);
  partition_kernel<<<(uint32_t)(((uint64_t)tiles + 1) * kSplitLanes + 255) / 256, 256, 0, ctx->stream>>>(
      ak, na, bk, nb, (uint32_t)tile, tiles, m.splits, m.counter);
  prof_end(ctx, pslot);
  LNT_HIP(hipGetLastError());
  *slot = prof_begin(ctx, kProfMerge, 16.0 * ((double)na + nb));
  merge_kernel<BT, ITEMS><<<tiles, BT, 0, ctx->stream>>>(
      ak, av, na, bk, bv, nb, op, ck, cv, m.splits, m.status,
      m.epoch, m.counter, m.total, tiles, capacity, ctx->tuning.narrow_keys ? 0 : 1);
  prof_end(ctx, *slot);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

// Steps per merge tile under the context's tuning (the tile launch_merge
// runs: merge_block x merge_items, 16 items when unset).
uint32_t merge_tile_steps(lnt_context ctx) {
  const int items = ctx->tuning.merge_items ? ctx->tuning.merge_items : 16;
  const int block = ctx->tuning.merge_block;
  switch (block * 100 + items) {
    case 25608: case 25612: case 25616: case 51208: case 51212: case 102408: case 102412: case 102416:
      return (uint32_t)(block * items);
    default:
      return 512 * 16;
  }
}

// Runs the merge with the given scratch; *nc_host = nnz(C).
int merge_with(lnt_context ctx, const uint64_t *ak, const double *av, uint32_t na, const uint64_t *bk,
               const double *bv, uint32_t nb, int op, uint64_t *ck, double *cv, uint64_t capacity,
               const MergeScratch &m, uint32_t *nc_host) {
  int rc, slot = -1;
#define LNT_MERGE(BT, IT) launch_merge<BT, IT>(ctx, ak, av, na, bk, bv, nb, op, ck, cv, capacity, m, &slot)
  // Measured on 2 x 50M-key merges: 512 x 16 (two tiles per CU).
  const int items = ctx->tuning.merge_items ? ctx->tuning.merge_items : 16;
  switch (ctx->tuning.merge_block * 100 + items) {
    case 25608: rc = LNT_MERGE(256, 8); break;
    case 25612: rc = LNT_MERGE(256, 12); break;
    case 25616: rc = LNT_MERGE(256, 16); break;
    case 51208: rc = LNT_MERGE(512, 8); break;
    case 51212: rc = LNT_MERGE(512, 12); break;
    case 102408: rc = LNT_MERGE(1024, 8); break;
    case 102412: rc = LNT_MERGE(1024, 12); break;
    case 102416: rc = LNT_MERGE(1024, 16); break;
    default: rc = LNT_MERGE(512, 16);
  }
#undef LNT_MERGE
  if (rc != LNT_OK) return rc;
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 1, m.total, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  *nc_host = ctx->pinned[1];
  prof_set_bytes(ctx, slot, 16.0 * ((double)na + nb + *nc_host));
  if (*nc_host > capacity)
    return fail(LNT_ERR_NOT_SORTED, "merge inputs are not sorted with unique keys (output overflow)");
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
  const uint32_t tiles = (uint32_t)(((uint64_t)na + nb + merge_tile_steps(ctx) - 1) / merge_tile_steps(ctx));
  LNT_TRY(ensure_status(ctx, tiles));
  LNT_TRY(ensure(ctx->aux, ((size_t)tiles + 1) * sizeof(uint32_t)));
  MergeScratch m;
  m.status = (uint64_t *)ctx->status.ptr;
  m.splits = (uint32_t *)ctx->aux.ptr;
  m.counter = small + kSmallCounters + 15;
  m.total = small + kSmallFlags + 1;
  LNT_TRY(next_epoch(ctx, &m.epoch));
  return merge_with(ctx, ak, av, na, bk, bv, nb, op, ck, cv, capacity, m, nc_host);
}

// ------------------------------------------- merge with no allocation
// A file-scope device array holds the scratch of lnt_merge_into_noalloc:
// kNoAllocTiles look-back words, then the merge-path splits (kNoAllocTiles + 1
// 32-bit words), the tile counter and the output total.  Every GPU the
// library runs on holds its own copy until the process ends (512 KiB); calls
// on one GPU share it.  Host code never passes the array's name to a kernel:
// it uses the address hipGetSymbolAddress gives for the device, kept per
// device by noalloc_scratch_init at context creation.
constexpr uint32_t kNoAllocTiles = 43000;
constexpr size_t kNoAllocWords = kNoAllocTiles + (kNoAllocTiles + 1 + 1) / 2 + 2;
static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

namespace {
constexpr int kMaxDevices = 64;
void *g_noalloc_ptr[kMaxDevices];
std::mutex g_noalloc_mutex;

__global__ void clear_words_kernel(uint64_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}
}  // namespace

// Looks up (once per device, with that device current) the address of this
// device's scratch array and keeps it in the context.
int noalloc_scratch_init(lnt_context ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(LNT_ERR_INVALID_ARGUMENT, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_ptr[ctx->device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(LNT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[ctx->device] = ptr;
  }
  ctx->noalloc_scratch = g_noalloc_ptr[ctx->device];
  return LNT_OK;
}

// The merge of lnt_merge_into with its scratch in the static array: no
// allocation of any kind.  The look-back words of the tiles this call uses
// are cleared by a kernel first, so a word left by an earlier call never
// reads as published; epoch 1 then serves every call.
int merge_noalloc_device(lnt_context ctx, const uint64_t *ak, const double *av, uint32_t na, const uint64_t *bk,
                         const double *bv, uint32_t nb, int op, uint64_t *ck, double *cv, uint64_t capacity,
                         uint32_t *nc_host) {
  if ((uint64_t)na + nb == 0) {
    *nc_host = 0;
    return LNT_OK;
  }
  if (!ctx->noalloc_scratch) return fail(LNT_ERR_DEVICE, "internal: the no-allocation scratch is not set up");
  const uint64_t tile = merge_tile_steps(ctx);
  const uint64_t tiles = ((uint64_t)na + nb + tile - 1) / tile;
  if (tiles > kNoAllocTiles)
    return fail(LNT_ERR_TOO_LARGE, "na + nb = " + std::to_string((uint64_t)na + nb) +
                                       " exceeds the no-allocation merge's limit of " +
                                       std::to_string(kNoAllocTiles * tile) + " (" + std::to_string(kNoAllocTiles) +
                                       " tiles of " + std::to_string(tile) + " steps)");
  uint64_t *base = (uint64_t *)ctx->noalloc_scratch;
  MergeScratch m;
  m.status = base;
  m.splits = (uint32_t *)(base + kNoAllocTiles);
  m.counter = m.splits + kNoAllocTiles + 1;
  m.total = m.counter + 1;
  m.epoch = 1;
  clear_words_kernel<<<(uint32_t)std::min<uint64_t>((tiles + 255) / 256, 1024), 256, 0, ctx->stream>>>(
      m.status, (uint32_t)tiles);
  LNT_HIP(hipGetLastError());
  return merge_with(ctx, ak, av, na, bk, bv, nb, op, ck, cv, capacity, m, nc_host);
}

// ------------------------------------------------- merges of 2^32 and more
namespace {
// The merge-path point at diagonal `diag` of the whole arrays (64-bit), moved
// one step on when A's last key before it equals B's next key, so that the pair
// stays in one chunk: out[0] = A entries, out[1] = B entries before the point.
__global__ void split_point_kernel(const uint64_t *__restrict__ a, uint64_t na, const uint64_t *__restrict__ b,
                                   uint64_t nb, uint64_t diag, uint64_t *out) {
  uint64_t lo = diag > nb ? diag - nb : 0, hi = diag < na ? diag : na;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] <= b[diag - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  uint64_t j = diag - lo;
  if (lo > 0 && j < nb && a[lo - 1] == b[j]) ++j;
  out[0] = lo;
  out[1] = j;
}
}  // namespace

// A merge whose inputs hold 2^32 entries or more together: the merge path is
// cut into chunks of at most 2^31 steps (plus the one step that keeps an
// equal pair together), each merged by merge_device into C at the running
// output count.
int merge_device_large(lnt_context ctx, const uint64_t *ak, const double *av, uint64_t na, const uint64_t *bk,
                       const double *bv, uint64_t nb, int op, uint64_t *ck, double *cv, uint64_t capacity,
                       uint64_t *nc_host) {
  constexpr uint64_t kChunkSteps = uint64_t(1) << 31;
  const uint64_t total = na + nb;
  LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
  uint64_t *pt = (uint64_t *)((uint32_t *)ctx->small.ptr + kSmallFlags + 6);
  uint64_t i = 0, j = 0, out = 0;
  while (i + j < total) {
    uint64_t ie = na, je = nb;
    if (i + j + kChunkSteps < total) {
      split_point_kernel<<<1, 1, 0, ctx->stream>>>(ak, na, bk, nb, i + j + kChunkSteps, pt);
      LNT_HIP(hipGetLastError());
      LNT_HIP(hipMemcpyAsync(ctx->pinned + 12, pt, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
      LNT_HIP(hipStreamSynchronize(ctx->stream));
      uint64_t p[2];
      memcpy(p, ctx->pinned + 12, sizeof(p));
      ie = p[0];
      je = p[1];
    }
    uint32_t count = 0;
    LNT_TRY(merge_device(ctx, ak + i, av + i, (uint32_t)(ie - i), bk + j, bv + j, (uint32_t)(je - j), op, ck + out,
                         cv + out, capacity - out, &count));
    out += count;
    i = ie;
    j = je;
  }
  *nc_host = out;
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
