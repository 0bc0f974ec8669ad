// sort.hip -- LSD onesweep radix sort of (uint64 key, double value) pairs for
// gfx950, with the relinearisation of change_sort_order fused into the first
// read of the keys.
//
// Replaces the reference's serial in-place MSD radix
// (multi_index_array/LibMIARadix.h:119-180, 444-484) and RadixShuffle
// (LibMIARadix.h:705-927).  Keys are unique, so the sorted order and the
// permuted values are fully determined and the result is bit-identical to the
// reference whatever the algorithm.
//
// Plan for b = ceil(log2(bound)) key bits: P = ceil(b/8) passes of 8-bit digits.
//   1. histogram kernel: one read of the keys (relinearised on the fly),
//      P digit histograms at once, key-range check;
//   2. P onesweep passes: each 256-thread block takes the next tile (dynamic
//      tile id) of 256*ITEMS pairs, ranks digits inside each wave with ballot
//      matching (stable), publishes its digit counts, obtains the global digit
//      offsets by decoupled look-back, stages the tile in LDS in digit order and
//      writes it out so that consecutive lanes store consecutive addresses.
//   Even P ping-pongs caller buffer -> scratch -> caller; odd P goes through two
//   scratch buffers so the result still lands in the caller's arrays.
#include <algorithm>
#include <atomic>
#include <vector>

#include "lnt_internal.h"

namespace lnt {

namespace {

constexpr int kBlock = 256;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = 8;

__global__ void fill_u32_kernel(uint32_t *p, uint32_t value, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = value;
}

__global__ void fill_u64_kernel(uint64_t *p, uint64_t value, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = value;
}

// ------------------------------------------------------------ histogram ---
struct DigitPlan {
  int passes;
  uint32_t shift[kMaxPasses];
  uint32_t bits[kMaxPasses];
};

template <bool RELIN>
__global__ __launch_bounds__(kBlock) void histogram_kernel(
    const uint64_t *__restrict__ keys, uint32_t n, DigitPlan plan, uint64_t bound,
    uint32_t *__restrict__ part, uint32_t *__restrict__ counters,
    uint32_t *__restrict__ hist, uint32_t *__restrict__ err, uint32_t tag, RelinDev *__restrict__ relin_out,
    const RelinDev relin) {
  const int passes = plan.passes;
  __shared__ uint32_t h[kMaxPasses * kRadix];
  const uint32_t tid = threadIdx.x;
  for (int i = tid; i < passes * kRadix; i += kBlock) h[i] = 0;
  if (blockIdx.x == 0) {
    // The radix passes read the table from device memory: this launch, the
    // first of every relinearising sort, stores it there.
    if (RELIN && tid == 0) *relin_out = relin;
    if (tid < 16) counters[tid] = 0;
    // histogram_reduce_kernel adds its row groups' sums into hist.
    for (int i = tid; i < passes * kRadix; i += kBlock) hist[i] = 0;
  }
  __syncthreads();

  constexpr int kUnroll = 8;
  bool bad = false;
  const uint64_t chunk = (uint64_t)kBlock * kUnroll;
  const uint64_t stride = (uint64_t)gridDim.x * chunk;
  // Unconditional loads of a clamped index: a guarded load compiles to a
  // branch that waits for it, which would serialise the kUnroll loads.
  auto load = [&](uint64_t (&k)[kUnroll], uint64_t base) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const uint64_t i = base + (uint64_t)u * kBlock + tid;
      k[u] = keys[i < n ? i : n - 1];
    }
  };
  auto count = [&](uint64_t (&k)[kUnroll], uint64_t base) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const uint64_t i = base + (uint64_t)u * kBlock + tid;
      bad |= i < n && k[u] >= bound;
    }
    if (RELIN) relin_keys<kUnroll>(&relin, k);
    // Consecutive lanes hold consecutive keys; inputs are often sorted or
    // nearly so.  A wave whose lanes all hold the same digit adds once
    // (sorted inputs make such waves common): no same-address LDS atomics
    // serialise inside a wave; otherwise every lane adds its own key.
    // The pass loop is unrolled to its maximum so each pass's shift and mask
    // are read once per kernel into scalar registers, not once per key.
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const uint64_t i = base + (uint64_t)u * kBlock + tid;
      const bool valid = i < n;
      const uint64_t key = k[u];
      const uint64_t vmask = __ballot(valid);
#pragma unroll
      for (int p = 0; p < kMaxPasses; ++p) {
        if (p < passes) {
          const uint32_t d = (uint32_t)(key >> plan.shift[p]) & ((1u << plan.bits[p]) - 1);
          const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
          if (__ballot(valid && d == d0) == vmask) {
            if ((tid & 63) == 0) atomicAdd(&h[p * kRadix + d0], (uint32_t)__popcll(vmask));
          } else if (valid) {
            atomicAdd(&h[p * kRadix + d], 1u);
          }
        }
      }
    }
  };
  // Two register buffers in turn: the next chunk's loads are issued before
  // the current chunk is counted, so a workgroup always has a chunk in flight.
  // The prefetches are unconditional (past the end they read key n-1): a
  // load on only one path makes the compiler wait for the older buffer's
  // loads as if the newer ones had not been issued.
  uint64_t k0[kUnroll], k1[kUnroll];
  uint64_t b0 = (uint64_t)blockIdx.x * chunk;
  if (b0 >= n) b0 = n;  // no chunk: the loop below does not run
  load(k0, b0 < n ? b0 : 0);
  while (b0 < n) {
    const uint64_t b1 = b0 + stride;
    load(k1, b1);
    count(k0, b0);
    if (b1 >= n) break;
    b0 = b1 + stride;
    load(k0, b0);
    count(k1, b1);
  }
  // The flag holds the tag of the last sort that saw a bad key: no clearing
  // launch is needed, and a stale value never equals a later sort's tag.
  if (bad) __hip_atomic_store(err, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  // Per-block partial histograms (no contended device-wide atomics: the
  // counters of all blocks would otherwise serialise across the XCDs' L2s).
  uint32_t *mine = part + (size_t)blockIdx.x * passes * kRadix;
  for (int i = tid; i < passes * kRadix; i += kBlock) mine[i] = h[i];
}

// hist[i] = sum over blocks of part[b][i].  Workgroup (x, y) covers 64
// consecutive (pass, digit) slots of rows [y*rows, (y+1)*rows); its four waves
// each sum every fourth row of that range with coalesced, unrolled loads, add
// up in LDS, and wave 0 adds the group's sums into hist (zeroed by
// histogram_kernel).  Splitting the rows over gridDim.y keeps each wave at one
// or two rounds of loads instead of a long serial walk over all rows.
__global__ __launch_bounds__(256) void histogram_reduce_kernel(const uint32_t *__restrict__ part,
                                                               uint32_t blocks, uint32_t width,
                                                               uint32_t rows,
                                                               uint32_t *__restrict__ hist) {
  __shared__ uint32_t s_sum[4][64];
  const uint32_t lane = threadIdx.x & 63, group = threadIdx.x >> 6;
  const uint32_t slot = blockIdx.x * 64 + lane;
  const uint32_t r0 = blockIdx.y * rows;
  const uint32_t r1 = min(blocks, r0 + rows);
  uint32_t acc = 0;
  if (slot < width) {
    uint32_t b = r0 + group;
    for (; b + 12 < r1; b += 16) {
      uint32_t x0 = part[(size_t)b * width + slot];
      uint32_t x1 = part[(size_t)(b + 4) * width + slot];
      uint32_t x2 = part[(size_t)(b + 8) * width + slot];
      uint32_t x3 = part[(size_t)(b + 12) * width + slot];
      acc += x0 + x1 + x2 + x3;
    }
    for (; b < r1; b += 4) acc += part[(size_t)b * width + slot];
  }
  s_sum[group][lane] = acc;
  __syncthreads();
  if (group == 0 && slot < width)
    atomicAdd(&hist[slot], s_sum[0][lane] + s_sum[1][lane] + s_sum[2][lane] + s_sum[3][lane]);
}

// -------------------------------------------------------------- onesweep ---
// KIn/KOut: key word read/written by this pass.  When the key space fits in
// 32 bits the passes between the first and the last move 32-bit keys (24 B per
// pair instead of 32 B); the first pass narrows, the last widens.
template <int BT, int ITEMS, bool RELIN, typename KIn, typename KOut>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4))) void onesweep_kernel(
    const KIn *__restrict__ kin, const double *__restrict__ vin,
    KOut *__restrict__ kout, double *__restrict__ vout, uint32_t n,
    uint32_t shift, uint32_t dbits, const uint32_t *__restrict__ pass_hist,
    uint64_t *__restrict__ status, uint32_t epoch,
    uint32_t *__restrict__ tile_counter, const uint32_t *__restrict__ stop, uint32_t stop_tag,
    const RelinDev *__restrict__ relin) {
  static_assert(BT >= kRadix && BT % kWave == 0, "one thread per digit at least");
  // stop: the histogram pass's key-range flag when the host checks it (set
  // to stop_tag on a bad key).  The flag is final before this pass starts, so
  // every workgroup of every pass makes the same choice: a rejected input is
  // left as it was.
  if (stop != nullptr && *stop == stop_tag) return;
  constexpr int kTile = BT * ITEMS;
  constexpr int kWaves = BT / kWave;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wcount[kWaves][kRadix];
  __shared__ uint32_t s_lstart[kRadix];
  __shared__ uint32_t s_gofs[kRadix];
  __shared__ uint32_t s_scan[kWaves];
  __shared__ uint64_t s_stage[kTile];

  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool digit_thread = tid < kRadix;  // one thread per digit for the per-digit steps
  if (tid == 0)
    s_tile = __hip_atomic_fetch_add(tile_counter, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t i = tid; i < (uint32_t)(kWaves * kRadix); i += BT) (&s_wcount[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t tile_base = (uint64_t)tile * kTile;
  const uint32_t tile_n = (uint32_t)min<uint64_t>(kTile, n - tile_base);
  const uint64_t wave_base = tile_base + (uint64_t)wave * (kWave * ITEMS);
  const uint32_t dmask = (1u << dbits) - 1;

  uint64_t k[ITEMS];
  double v[ITEMS];
  // Unconditional loads of a clamped index, so all 2*ITEMS loads are in
  // flight together (a guarded load becomes a branch that waits for it).
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint64_t pos = wave_base + (uint64_t)i * kWave + lane;
    uint64_t at = pos < n ? pos : n - 1;
    k[i] = (uint64_t)kin[at];
    v[i] = vin[at];
  }
  if (RELIN) relin_keys<ITEMS>(relin, k);

  // Stable in-wave ranking: the lanes holding the same digit are the lanes
  // that agree with this one on every digit bit.  Per bit: m = the bit
  // sign-extended (0 or ~0), its ballot, and the lanes that differ,
  // ballot ^ m, or-ed into a mismatch mask -- one v_bitop3 per half (truth
  // table 0xF6 = src0 | (src1 ^ src2)).  Lanes past the end start as
  // mismatches.  Bits above dbits are zero in every lane, so a fixed 8-bit
  // loop serves every digit width.
  uint32_t local_rank[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint64_t pos = wave_base + (uint64_t)i * kWave + lane;
    const bool valid = pos < n;
    uint32_t d = (uint32_t)(k[i] >> shift) & dmask;
    const uint64_t vb = __ballot(valid);
    uint32_t x_lo = ~(uint32_t)vb, x_hi = ~(uint32_t)(vb >> 32);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)d, b, 1);  // 0 or ~0
      const uint64_t bb = __ballot(m != 0);
      x_lo = __builtin_amdgcn_bitop3_b32(x_lo, (uint32_t)bb, m, 0xF6);
      x_hi = __builtin_amdgcn_bitop3_b32(x_hi, (uint32_t)(bb >> 32), m, 0xF6);
    }
    const uint64_t peers = ~(((uint64_t)x_hi << 32) | x_lo);
    uint32_t r = 0;
    if (valid) {
      // Every lane of a group reads the count before any writes it, and all
      // of them store the same new count: no leader election, no branch.
      const uint32_t cnt = s_wcount[wave][d];
      r = cnt + mask_rank(peers);
      s_wcount[wave][d] = cnt + (uint32_t)__popcll(peers);
    }
    local_rank[i] = r;
  }
  __syncthreads();

  // Per digit (one thread each): prefix over waves, then over digits.
  uint32_t count = 0;
  if (digit_thread) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      uint32_t c = s_wcount[w][tid];
      s_wcount[w][tid] = count;
      count += c;
    }
  }
  uint32_t total;
  const uint32_t lstart = block_exclusive_sum_bt<BT>(count, s_scan, &total);
  uint32_t hist0 = 0;
  if (tile == 0 && digit_thread) hist0 = pass_hist[tid];
  uint32_t all;
  const uint32_t excl0 = tile == 0 ? block_exclusive_sum_bt<BT>(hist0, s_scan, &all) : 0;
  // The tile's digit counts are published at once; the look-back for its
  // global offsets runs after this thread has staged its keys (staging needs
  // only the tile-local offsets).
  uint64_t *my_status = status + (uint64_t)tile * kRadix + tid;
  if (digit_thread) {
    s_lstart[tid] = lstart;
    status_store(my_status, tile == 0 ? status_word(kFlagInclusive, epoch, excl0 + count)
                                      : status_word(kFlagAggregate, epoch, count));
  }
  __syncthreads();

  // Keys: stage in digit order, then store.
  uint32_t lpos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint64_t pos = wave_base + (uint64_t)i * kWave + lane;
    uint32_t d = (uint32_t)(k[i] >> shift) & dmask;
    lpos[i] = 0xFFFFFFFFu;
    if (pos < n) {
      lpos[i] = s_lstart[d] + s_wcount[wave][d] + local_rank[i];
      s_stage[lpos[i]] = k[i];
    }
  }
  if (digit_thread) {
    uint32_t excl = excl0;
    if (tile > 0) {
      excl = lookback_batched<4>(status, tile, kRadix, tid, epoch);
      status_store(my_status, status_word(kFlagInclusive, epoch, excl + count));
    }
    s_gofs[tid] = excl - lstart;
  }
  __syncthreads();
  // Destinations stay below n whenever the digit counts of the histogram
  // pass agree with the keys seen here; the bound keeps a malformed input
  // (validation off) from writing outside the output.
  uint32_t dst[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint32_t p = (uint32_t)i * BT + tid;
    dst[i] = 0xFFFFFFFFu;
    if (p < tile_n) {
      uint64_t key = s_stage[p];
      uint32_t d = (uint32_t)(key >> shift) & dmask;
      dst[i] = s_gofs[d] + p;
      if (dst[i] < n) kout[dst[i]] = (KOut)key;
    }
  }
  __syncthreads();
  // Values: same permutation.
  double *s_vstage = reinterpret_cast<double *>(s_stage);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if (lpos[i] != 0xFFFFFFFFu) s_vstage[lpos[i]] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint32_t p = (uint32_t)i * BT + tid;
    if (p < tile_n && dst[i] < n) vout[dst[i]] = s_vstage[p];
  }
}

__global__ void relin_kernel(uint64_t *keys, uint32_t n, const RelinDev relin) {
  constexpr int kPer = 4;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; base < n; base += stride * kPer) {
    uint64_t k[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      uint64_t i = base + (uint64_t)u * stride;
      k[u] = keys[i < n ? i : n - 1];
    }
    relin_keys<kPer>(&relin, k);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      uint64_t i = base + (uint64_t)u * stride;
      if (i < n) keys[i] = k[u];
    }
  }
}

enum KeyWidth { kW64 = 0, kW32 = 1 };

template <int BT, int ITEMS, bool RELIN, typename KIn, typename KOut>
void launch_one(lnt_context ctx, uint32_t tiles, const void *kin, const double *vin, void *kout, double *vout,
                uint32_t n, uint32_t shift, uint32_t dbits, const uint32_t *pass_hist, uint64_t *status,
                uint32_t epoch, uint32_t *counter, const RelinDev *r, const uint32_t *stop, uint32_t tag) {
  onesweep_kernel<BT, ITEMS, RELIN, KIn, KOut><<<tiles, BT, 0, ctx->stream>>>(
      (const KIn *)kin, vin, (KOut *)kout, vout, n, shift, dbits, pass_hist, status, epoch, counter, stop, tag, r);
}

template <int BT, int ITEMS>
int launch_pass(lnt_context ctx, bool relin, int win, int wout, const void *kin,
                const double *vin, void *kout, double *vout, uint32_t n,
                uint32_t shift, uint32_t dbits, const uint32_t *pass_hist,
                uint32_t *counter, const RelinDev *r, const uint32_t *stop, uint32_t tag) {
  constexpr uint32_t tile = BT * ITEMS;
  uint32_t tiles = (uint32_t)(((uint64_t)n + tile - 1) / tile);
  LNT_TRY(ensure_status(ctx, (size_t)tiles * kRadix));
  uint32_t epoch;
  LNT_TRY(next_epoch(ctx, &epoch));
  uint64_t *status = (uint64_t *)ctx->status.ptr;
  if (ctx->capturing) {
    // A recorded graph replays this launch with the same epoch: clear the
    // words the pass uses so that a replay never sees its predecessor's.
    size_t words = (size_t)tiles * kRadix;
    fill_u64_kernel<<<(uint32_t)std::min<size_t>((words + 255) / 256, 4096), 256, 0, ctx->stream>>>(status, 0, words);
    LNT_HIP(hipGetLastError());
  }
  // Roofline basis: the fixed model of 32 B per pair per pass (8 B key + 8 B
  // value, read and written), whatever key width this pass actually moves.
  int slot = prof_begin(ctx, kProfOnesweep, 32.0 * n);
  const int sel = (relin ? 4 : 0) + win * 2 + wout;
  switch (sel) {
    case 0: launch_one<BT, ITEMS, false, uint64_t, uint64_t>(ctx, tiles, kin, vin, kout, vout, n, shift, dbits, pass_hist, status, epoch, counter, r, stop, tag); break;
    case 1: launch_one<BT, ITEMS, false, uint64_t, uint32_t>(ctx, tiles, kin, vin, kout, vout, n, shift, dbits, pass_hist, status, epoch, counter, r, stop, tag); break;
    case 2: launch_one<BT, ITEMS, false, uint32_t, uint64_t>(ctx, tiles, kin, vin, kout, vout, n, shift, dbits, pass_hist, status, epoch, counter, r, stop, tag); break;
    case 3: launch_one<BT, ITEMS, false, uint32_t, uint32_t>(ctx, tiles, kin, vin, kout, vout, n, shift, dbits, pass_hist, status, epoch, counter, r, stop, tag); break;
    case 4: launch_one<BT, ITEMS, true, uint64_t, uint64_t>(ctx, tiles, kin, vin, kout, vout, n, shift, dbits, pass_hist, status, epoch, counter, r, stop, tag); break;
    case 5: launch_one<BT, ITEMS, true, uint64_t, uint32_t>(ctx, tiles, kin, vin, kout, vout, n, shift, dbits, pass_hist, status, epoch, counter, r, stop, tag); break;
    default: return fail(LNT_ERR_INVALID_ARGUMENT, "internal: unsupported key widths");
  }
  prof_end(ctx, slot);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

}  // namespace

int launch_fill_u32(lnt_context ctx, uint32_t *p, uint32_t value, size_t n) {
  if (n == 0) return LNT_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  fill_u32_kernel<<<(uint32_t)blocks, 256, 0, ctx->stream>>>(p, value, n);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

int ensure_status(lnt_context ctx, size_t words) {
  size_t bytes = words * sizeof(uint64_t);
  if (ctx->status.bytes >= bytes) return LNT_OK;
  // Over-allocate so that slightly larger sorts do not reallocate.
  LNT_TRY(ensure(ctx->status, bytes + bytes / 4));
  size_t all = ctx->status.bytes / sizeof(uint64_t);
  fill_u64_kernel<<<4096, 256, 0, ctx->stream>>>((uint64_t *)ctx->status.ptr, 0, all);
  LNT_HIP(hipGetLastError());
  ctx->epoch = 0;  // fresh words carry epoch 0, which is never current
  return LNT_OK;
}

int next_epoch(lnt_context ctx, uint32_t *epoch) {
  ctx->epoch = (ctx->epoch + 1) & kEpochMask;
  if (ctx->epoch == 0) {
    size_t all = ctx->status.bytes / sizeof(uint64_t);
    fill_u64_kernel<<<4096, 256, 0, ctx->stream>>>((uint64_t *)ctx->status.ptr, 0, all);
    LNT_HIP(hipGetLastError());
    ctx->epoch = 1;
  }
  *epoch = ctx->epoch;
  return LNT_OK;
}

// Tags of the sorts' key-range flags: process-wide, so that a flag word left
// by any earlier sort (of any context, in memory since reused) never matches.
static uint32_t next_check_tag() {
  static std::atomic<uint32_t> counter{0};
  uint32_t t;
  do t = counter.fetch_add(1, std::memory_order_relaxed) + 1;
  while (t == 0);
  return t;
}

int relin_device(lnt_context ctx, uint64_t *keys, uint32_t n, const Relin &relin) {
  if (n == 0 || !relin.active) return LNT_OK;
  uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)n + 255) / 256, 8192);
  int slot = prof_begin(ctx, kProfRelin, 16.0 * n);
  relin_kernel<<<blocks, 256, 0, ctx->stream>>>(keys, n, relin_for_device(relin));
  prof_end(ctx, slot);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

// Sorts (keys, vals) in place.  bound: keys must lie in [0, bound) before the
// optional relinearisation (which maps [0,bound) onto itself).
uint32_t sorted_low_bits(int order, const int64_t *dims, const int32_t *cur_seq, const int32_t *new_seq) {
  int pos[LNT_MAX_ORDER];
  for (int t = 0; t < order; ++t) pos[cur_seq[t]] = t;
  uint64_t p = 1;
  for (int j = 0, last = -1; j < order && pos[new_seq[j]] > last; ++j) {
    last = pos[new_seq[j]];
    p *= (uint64_t)dims[new_seq[j]];
  }
  uint32_t b = 0;
  while (b < 63 && !((p >> b) & 1)) ++b;
  return b;
}

int sort_device(lnt_context ctx, uint64_t *keys, double *vals, uint32_t n,
                uint64_t bound, const Relin *relin, uint32_t sorted_low) {
  if (n == 0) return LNT_OK;
  const bool do_relin = relin && relin->active;
  Relin r{};
  if (do_relin) r = *relin;
  // The table travels by value as the histogram kernel's last argument
  // (captured with the launch; last, so that the other arguments share no
  // cache line with it), which stores it for the radix passes.
  const RelinDev rtab = do_relin ? relin_for_device(r) : RelinDev{};
  RelinDev *rdev = nullptr;
  if (do_relin) {
    LNT_TRY(ensure(ctx->relin_dev, sizeof(RelinDev)));
    rdev = (RelinDev *)ctx->relin_dev.ptr;
  }
  uint32_t bits = 0;
  while (bits < 64 && (uint64_t(1) << bits) < bound) ++bits;
  // Digit plan over the bits above the already ordered low bits:
  // ceil(span/digit_bits) passes (at least two so that no pass runs in
  // place), digit widths balanced, least significant digit first.
  DigitPlan plan{};
  const uint32_t lo = std::min(sorted_low, bits), span = bits - lo;
  const uint32_t dmax = (uint32_t)std::min(std::max(ctx->tuning.digit_bits, 4), kRadixBits);
  int passes = (int)((span + dmax - 1) / dmax);
  if (passes > kMaxPasses) passes = kMaxPasses;  // 64-bit keys with narrow digits: widen
  if (passes == 1) passes = 2;
  if (n == 1 || span == 0) passes = 0;
  plan.passes = passes;
  uint32_t shift = lo;
  for (int p = 0; p < passes; ++p) {
    uint32_t w = span / passes + ((uint32_t)p < span % passes ? 1 : 0);
    plan.shift[p] = shift;
    plan.bits[p] = w;
    shift += w;
  }
  LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
  uint32_t *small = (uint32_t *)ctx->small.ptr;
  uint32_t *hist = small + kSmallHist;
  uint32_t *counters = small + kSmallCounters;
  uint32_t *err = small + kSmallFlags;
  const uint32_t tag = next_check_tag();

  uint32_t hblocks = (uint32_t)ctx->tuning.hist_blocks;
  uint32_t need = (uint32_t)(((uint64_t)n + kBlock * 8 - 1) / (kBlock * 8));
  if (hblocks > need) hblocks = need;
  const uint32_t width = (uint32_t)(passes > 0 ? passes : 1) * kRadix;
  LNT_TRY(ensure(ctx->hist_part, (size_t)kMaxHistBlocks * kMaxPasses * kRadix * sizeof(uint32_t)));
  uint32_t *part = (uint32_t *)ctx->hist_part.ptr;
  int slot = prof_begin(ctx, kProfHistogram, 8.0 * n);
  if (do_relin)
    histogram_kernel<true><<<hblocks, kBlock, 0, ctx->stream>>>(keys, n, plan, bound, part, counters, hist, err, tag, rdev, rtab);
  else
    histogram_kernel<false><<<hblocks, kBlock, 0, ctx->stream>>>(keys, n, plan, bound, part, counters, hist, err, tag, rdev, rtab);
  prof_end(ctx, slot);
  LNT_HIP(hipGetLastError());
  if (passes > 0) {
    constexpr uint32_t kRows = 64;  // partial rows per reduce workgroup
    histogram_reduce_kernel<<<dim3((width + 63) / 64, (hblocks + kRows - 1) / kRows), 256, 0, ctx->stream>>>(
        part, hblocks, width, kRows, hist);
    LNT_HIP(hipGetLastError());
  }

  if (passes > 0) {  // ping-pong buffers, allocated before the flag copy is queued
    LNT_TRY(ensure(ctx->alt_keys, (size_t)n * 8));
    LNT_TRY(ensure(ctx->alt_vals, (size_t)n * 8));
    if (passes & 1) {
      LNT_TRY(ensure(ctx->alt2_keys, (size_t)n * 8));
      LNT_TRY(ensure(ctx->alt2_vals, (size_t)n * 8));
    }
  }
  // Key-range check: the flag is copied out and, when there are passes, they
  // are queued before the host waits for the copy (an event, not the stream),
  // so the GPU goes on with them meanwhile; on a bad key every pass returns at
  // once (its stop argument) and the arrays are left as they were.
  const bool check = ctx->validate && !ctx->capturing;
  if (check) {
    LNT_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    if (passes == 0) {
      LNT_HIP(hipStreamSynchronize(ctx->stream));
    } else {
      if (!ctx->check_event) LNT_HIP(hipEventCreateWithFlags(&ctx->check_event, hipEventDisableTiming));
      LNT_HIP(hipEventRecord(ctx->check_event, ctx->stream));
    }
  }
  auto key_range_ok = [&]() -> int {
    if (check) {
      if (passes > 0) LNT_HIP(hipEventSynchronize(ctx->check_event));
      if (ctx->pinned[0] == tag) return fail(LNT_ERR_KEY_RANGE, "a key is negative or not below the dimensionality");
    }
    return LNT_OK;
  };
  if (passes == 0) {
    LNT_TRY(key_range_ok());
    if (do_relin) LNT_TRY(relin_device(ctx, keys, n, r));
    return LNT_OK;
  }

  uint64_t *bufk[3] = {keys, (uint64_t *)ctx->alt_keys.ptr, nullptr};
  double *bufv[3] = {vals, (double *)ctx->alt_vals.ptr, nullptr};
  if (passes & 1) {
    bufk[2] = (uint64_t *)ctx->alt2_keys.ptr;
    bufv[2] = (double *)ctx->alt2_vals.ptr;
  }
  // Buffers: even P: 0->1->0->...->0; odd P: 0->1->2->1->...->0.  No pass
  // reads and writes the same buffer.  Keys of at most 32 bits travel as
  // 32-bit words between the first and the last pass.
  const bool narrow = bits <= 32 && ctx->tuning.narrow_keys;
  int src = 0;
  for (int p = 0; p < passes; ++p) {
    int dst;
    if (p == passes - 1) dst = 0;
    else if (passes & 1) dst = (src == 1) ? 2 : 1;
    else dst = (src == 0) ? 1 : 0;
    const bool rl = do_relin && p == 0;
    const int win = (narrow && p > 0) ? kW32 : kW64;
    const int wout = (narrow && p < passes - 1) ? kW32 : kW64;
    int rc;
    const int cfg = ctx->tuning.sort_block == 1024  ? 200 + ctx->tuning.items_per_thread
                    : ctx->tuning.sort_block == 512 ? 100 + ctx->tuning.items_per_thread
                                                    : ctx->tuning.items_per_thread;
#define LNT_PASS(BT, IT)                                                                                     \
  launch_pass<BT, IT>(ctx, rl, win, wout, bufk[src], bufv[src], bufk[dst], bufv[dst], n, plan.shift[p], \
                      plan.bits[p], hist + p * kRadix, counters + p, rdev, check ? err : nullptr, tag)
    switch (cfg) {
      case 8: rc = LNT_PASS(256, 8); break;
      case 12: rc = LNT_PASS(256, 12); break;
      case 108: rc = LNT_PASS(512, 8); break;
      case 112: rc = LNT_PASS(512, 12); break;
      case 116: rc = LNT_PASS(512, 16); break;
      case 208: rc = LNT_PASS(1024, 8); break;
      case 212: rc = LNT_PASS(1024, 12); break;
      case 216: rc = LNT_PASS(1024, 12); break;  // 1024x16: 128 VGPRs with spills, slower
      default: rc = LNT_PASS(256, 16); break;
    }
#undef LNT_PASS
    if (rc != LNT_OK) {
      (void)key_range_ok();  // the copy must land before pinned[0] is reused
      return rc;
    }
    src = dst;
  }
  return key_range_ok();
}

namespace {
// pos[b] = first position of a key >= b << shift in a run partitioned by
// key >> shift (pos[nb] = n).
__global__ void bucket_bounds_kernel(const uint64_t *__restrict__ keys, uint64_t n, uint32_t shift, uint32_t nb,
                                     uint64_t *__restrict__ pos) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nb) return;
  uint64_t lo = 0, hi = n;
  if (b == nb) lo = n;
  const uint64_t t = (uint64_t)b << shift;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((keys[mid] >> shift) < (t >> shift)) lo = mid + 1;
    else hi = mid;
  }
  pos[b] = lo;
}

// Moves each entry of one partitioned run to its bucket's place in the
// output: position + delta[bucket].  Keys beyond the bound (unvalidated input)
// are clamped to the last bucket and writes past `total` are dropped, so that
// every access stays in bounds.
__global__ void bucket_gather_kernel(const uint64_t *__restrict__ k, const double *__restrict__ v, uint64_t n,
                                     uint32_t shift, uint32_t nb, const int64_t *__restrict__ delta,
                                     uint64_t total, uint64_t *__restrict__ ok, double *__restrict__ ov) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t key = k[i];
    const uint64_t b = (key >> shift) < nb ? (key >> shift) : nb - 1;
    const uint64_t d = (uint64_t)((int64_t)i + delta[b]);
    if (d < total) {
      ok[d] = key;
      ov[d] = v[i];
    }
  }
}

__global__ void key_bound_kernel(const uint64_t *__restrict__ k, uint64_t n, uint64_t bound, uint32_t *err) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) bad |= k[i] >= bound;
  if (bad) *err = 1;
}
}  // namespace

int relin_device_large(lnt_context ctx, uint64_t *keys, uint64_t n, const Relin &relin) {
  constexpr uint64_t kChunk = uint64_t(1) << 31;
  for (uint64_t s = 0; s < n; s += kChunk)
    LNT_TRY(relin_device(ctx, keys + s, (uint32_t)std::min(kChunk, n - s), relin));
  return LNT_OK;
}

// A sort of 2^32 entries or more, stable like sort_device:
//  1. each chunk of 2^31 entries is relinearised and stably ordered by the
//     top h = min(bits, 16) key bits only (sort_device with the low bits
//     declared ordered: two 8-bit passes);
//  2. the 2^h bucket boundaries of every chunk are found by binary search;
//  3. every chunk's buckets are gathered into bucket order (chunk order
//     within a bucket, so equal keys keep their input order);
//  4. consecutive buckets are grouped up to 2^31 entries and each group is
//     sorted on all bits by sort_device (a bucket of one key value needs no
//     sort: when h = bits every bucket is one value);
//  5. the result is copied back in place.
// Work memory: 16 bytes per entry for the gathered copy, plus sort_device's
// ping-pong buffers for one chunk or group.
int sort_device_large(lnt_context ctx, uint64_t *keys, double *vals, uint64_t n, uint64_t bound,
                      const Relin *relin) {
  constexpr uint64_t kChunk = uint64_t(1) << 31;
  uint32_t bits = 0;
  while (bits < 64 && (uint64_t(1) << bits) < bound) ++bits;
  const uint32_t h = std::min<uint32_t>(bits, 16), shift = bits - h, nb = 1u << h;
  const uint64_t chunks = (n + kChunk - 1) / kChunk;
  if (ctx->validate) {  // before any chunk moves, so that a bad key leaves the arrays as they were
    LNT_TRY(ensure(ctx->small, kSmallWords * sizeof(uint32_t)));
    uint32_t *err = (uint32_t *)ctx->small.ptr + kSmallFlags + 2;
    LNT_TRY(launch_fill_u32(ctx, err, 0, 1));
    key_bound_kernel<<<16384, 256, 0, ctx->stream>>>(keys, n, bound, err);
    LNT_HIP(hipGetLastError());
    LNT_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->pinned[0]) return fail(LNT_ERR_KEY_RANGE, "a key is negative or not below the dimensionality");
  }
  for (uint64_t c = 0; c < chunks; ++c) {
    const uint64_t s = c * kChunk;
    LNT_TRY(sort_device(ctx, keys + s, vals + s, (uint32_t)std::min(kChunk, n - s), bound, relin, shift));
  }
  Staging w;
  LNT_TRY(ensure(w.b[0], chunks * (nb + 1) * sizeof(uint64_t)));
  LNT_TRY(ensure(w.b[1], chunks * nb * sizeof(int64_t)));
  LNT_TRY(ensure(w.b[2], n * sizeof(uint64_t)));
  LNT_TRY(ensure(w.b[3], n * sizeof(double)));
  uint64_t *pos = (uint64_t *)w.b[0].ptr;
  for (uint64_t c = 0; c < chunks; ++c) {
    const uint64_t s = c * kChunk;
    bucket_bounds_kernel<<<(nb + 1 + 255) / 256, 256, 0, ctx->stream>>>(keys + s, std::min(kChunk, n - s), shift,
                                                                        nb, pos + c * (nb + 1));
    LNT_HIP(hipGetLastError());
  }
  std::vector<uint64_t> hp(chunks * (nb + 1));
  LNT_HIP(hipMemcpyAsync(hp.data(), pos, hp.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  // Bucket sizes, their places in the output, and each chunk's offsets.
  std::vector<uint64_t> base(nb + 1, 0);
  for (uint32_t b = 0; b < nb; ++b) {
    uint64_t t = 0;
    for (uint64_t c = 0; c < chunks; ++c) {
      const uint64_t lo = hp[c * (nb + 1) + b], hi = hp[c * (nb + 1) + b + 1];
      if (hi < lo) return fail(LNT_ERR_KEY_RANGE, "a key is negative or not below the dimensionality");
      t += hi - lo;
    }
    if (t >= (uint64_t(1) << 32) && shift > 0)
      return fail(LNT_ERR_TOO_LARGE, "2^32 or more keys share their top 16 bits");
    base[b + 1] = base[b] + t;
  }
  std::vector<int64_t> delta(chunks * nb);
  for (uint32_t b = 0; b < nb; ++b) {
    uint64_t at = base[b];
    for (uint64_t c = 0; c < chunks; ++c) {
      const uint64_t lo = hp[c * (nb + 1) + b], hi = hp[c * (nb + 1) + b + 1];
      delta[c * nb + b] = (int64_t)at - (int64_t)lo;
      at += hi - lo;
    }
  }
  int64_t *ddev = (int64_t *)w.b[1].ptr;
  LNT_HIP(hipMemcpyAsync(ddev, delta.data(), delta.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
  uint64_t *tk = (uint64_t *)w.b[2].ptr;
  double *tv = (double *)w.b[3].ptr;
  for (uint64_t c = 0; c < chunks; ++c) {
    const uint64_t s = c * kChunk;
    bucket_gather_kernel<<<16384, 256, 0, ctx->stream>>>(keys + s, vals + s, std::min(kChunk, n - s), shift, nb,
                                                         ddev + c * nb, n, tk, tv);
    LNT_HIP(hipGetLastError());
  }
  if (shift > 0) {
    for (uint32_t b = 0; b < nb;) {
      uint32_t e = b + 1;
      while (e < nb && base[e + 1] - base[b] <= kChunk) ++e;
      const uint64_t len = base[e] - base[b];
      if (len > 1) LNT_TRY(sort_device(ctx, tk + base[b], tv + base[b], (uint32_t)len, bound, nullptr, 0));
      b = e;
    }
  }
  LNT_HIP(hipMemcpyAsync(keys, tk, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
  LNT_HIP(hipMemcpyAsync(vals, tv, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  return LNT_OK;
}

}  // namespace lnt
