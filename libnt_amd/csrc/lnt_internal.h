// lnt_internal.h -- shared internals of the gfx950 sparse multi-index-array
// library: exact 64-bit division by invariant divisors, the relinearisation
// descriptor, decoupled look-back status words, the context, and error
// reporting.  Device code is written for CDNA4 (wave64) only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/lnt.h"

namespace lnt {

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors ---
void set_error(const std::string &msg);
int fail(int status, const std::string &msg);

#define LNT_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess)                                                    \
      return ::lnt::fail(LNT_ERR_DEVICE, std::string(#call) + ": " +        \
                                             hipGetErrorString(_e));         \
  } while (0)

#define LNT_TRY(expr)          \
  do {                         \
    int _s = (expr);           \
    if (_s != LNT_OK) return _s; \
  } while (0)

// ------------------------------------------------- exact unsigned division ---
// q = n / d for a divisor fixed per launch, by the multiply-high method of
// Granlund & Montgomery ("Division by invariant integers using
// multiplication", PLDI 1994, Fig. 4.1): with l = ceil(log2 d) and
// m = floor(2^64 (2^l - d) / d) + 1,
//   t = mulhi(m, n);  q = (t + ((n - t) >> s1)) >> s2,  s1 = min(l,1), s2 = max(l-1,0)
// is exact for every 64-bit n and every 1 <= d < 2^64.
struct Divider {
  uint64_t d;
  uint64_t magic;
  uint32_t s1, s2;
};

inline Divider make_divider(uint64_t d) {
  Divider v;
  v.d = d;
  uint32_t l = 0;
  while (l < 64 && (uint64_t(1) << l) < d) ++l;
  unsigned __int128 two_l = (unsigned __int128)1 << l;
  unsigned __int128 m = (((unsigned __int128)1 << 64) * (two_l - d)) / d + 1;
  v.magic = (uint64_t)m;
  v.s1 = l > 0 ? 1 : 0;
  v.s2 = l > 0 ? l - 1 : 0;
  return v;
}

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ inline uint64_t divide(uint64_t n, const Divider &v) {
  uint64_t t = mulhi64(v.magic, n);
  return (t + ((n - t) >> v.s1)) >> v.s2;
}

// ------------------------------------------------------ relinearisation ---
// New key from old key: peel the old-order digits from the most significant
// down (division by the old place values) and re-weight each by the place
// value of the same mode in the new order.  Closed form of
// sub2ind_reorder(ind2sub_reorder(key, cur), new).
struct Relin {
  int32_t order;
  int32_t active;  // 0: identity
  int32_t pow2;    // every dimension is a power of two: shifts and masks only
  int32_t pad;
  Divider place[LNT_MAX_ORDER];    // old place value of old position t
  uint64_t weight[LNT_MAX_ORDER];  // new place value of the mode at old position t
  uint8_t old_shift[LNT_MAX_ORDER], bits[LNT_MAX_ORDER], new_shift[LNT_MAX_ORDER];
};

// Device-resident form of a Relin: word-sized fields, read by every lane
// with uniform addresses (scalar loads) inside a loop over the runtime order.
// (A by-value kernel argument indexed at run time is copied to private
// memory in every lane.)
struct RelinDev {
  uint64_t magic[LNT_MAX_ORDER], d[LNT_MAX_ORDER], weight[LNT_MAX_ORDER];
  uint32_t s1[LNT_MAX_ORDER], s2[LNT_MAX_ORDER];
  uint32_t old_shift[LNT_MAX_ORDER], bits[LNT_MAX_ORDER], new_shift[LNT_MAX_ORDER];
  int32_t order, pow2;
};

inline RelinDev relin_for_device(const Relin &r) {
  RelinDev d{};
  for (int t = 0; t < LNT_MAX_ORDER; ++t) {
    d.magic[t] = r.place[t].magic;
    d.d[t] = r.place[t].d;
    d.s1[t] = r.place[t].s1;
    d.s2[t] = r.place[t].s2;
    d.weight[t] = r.weight[t];
    d.old_shift[t] = r.old_shift[t];
    d.bits[t] = r.bits[t];
    d.new_shift[t] = r.new_shift[t];
  }
  d.order = r.order;
  d.pow2 = r.pow2;
  return d;
}

// Relinearises N keys at once: modes outermost, so each mode's parameters
// are loaded once (uniform loads) and applied to all N keys.
template <int N>
__device__ __forceinline__ void relin_keys(const RelinDev *__restrict__ s, uint64_t (&k)[N]) {
  const int order = s->order;
  uint64_t out[N];
#pragma unroll
  for (int i = 0; i < N; ++i) out[i] = 0;
  if (s->pow2) {
    for (int t = 0; t < order; ++t) {
      const uint32_t os = s->old_shift[t], b = s->bits[t], ns = s->new_shift[t];
      const uint64_t mask = b >= 64 ? ~0ull : ((uint64_t(1) << b) - 1);
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] |= ((k[i] >> os) & mask) << ns;
    }
  } else {
    for (int t = order - 1; t >= 1; --t) {
      const uint64_t m = s->magic[t], d = s->d[t], w = s->weight[t];
      const uint32_t s1 = s->s1[t], s2 = s->s2[t];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const uint64_t hi = mulhi64(m, k[i]);
        const uint64_t q = (hi + ((k[i] - hi) >> s1)) >> s2;
        out[i] += q * w;
        k[i] -= q * d;
      }
    }
    const uint64_t w0 = s->weight[0];
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] += k[i] * w0;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) k[i] = out[i];
}

// Validates (order, dims, seqs) and fills a Relin.  *dimensionality receives
// prod(dims).  Returns an lnt_status.
int make_relin(int order, const int64_t *dims, const int32_t *cur_seq,
               const int32_t *new_seq, Relin *out, uint64_t *dimensionality);

// ----------------------------------------------- decoupled look-back word ---
// 64-bit status word per (tile, digit): [63:62] flag, [61:32] epoch, [31:0]
// value.  A word whose epoch differs from the current pass's is "not yet
// published", so the status array never needs clearing between passes.
constexpr uint64_t kFlagAggregate = 1;
constexpr uint64_t kFlagInclusive = 2;
constexpr uint32_t kEpochMask = 0x3FFFFFFFu;

__device__ inline uint64_t status_word(uint64_t flag, uint32_t epoch,
                                       uint32_t value) {
  return (flag << 62) | ((uint64_t)(epoch & kEpochMask) << 32) | value;
}

__device__ inline void status_store(uint64_t *p, uint64_t w) {
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t status_load(const uint64_t *p) {
  return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of this tile's value among all earlier tiles, by walking
// back until an inclusive word is found.  Tile 0 never calls this.
__device__ inline uint32_t lookback(const uint64_t *status, uint32_t tile,
                                    uint32_t stride, uint32_t slot,
                                    uint32_t epoch) {
  uint32_t excl = 0;
  int64_t t = (int64_t)tile - 1;
  while (t >= 0) {
    uint64_t w = status_load(status + (uint64_t)t * stride + slot);
    uint64_t flag = w >> 62;
    if (((uint32_t)(w >> 32) & kEpochMask) != (epoch & kEpochMask) || flag == 0) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += (uint32_t)w;
    if (flag == kFlagInclusive) break;
    --t;
  }
  return excl;
}

// Same walk with K status words of one slot loaded per round trip (the loads
// of one batch are independent, so a walk over many tiles that have only
// published their aggregates costs 1/K of the round trips).
template <int K>
__device__ inline uint32_t lookback_batched(const uint64_t *status, uint32_t tile,
                                            uint32_t stride, uint32_t slot,
                                            uint32_t epoch) {
  uint32_t excl = 0;
  int64_t t = (int64_t)tile - 1;
  // The nearest predecessor usually already holds its inclusive prefix:
  // read it alone first and batch only when the walk has to go further
  // (each extra word read is 8 B per digit of every tile).
  {
    uint64_t w = status_load(status + (uint64_t)t * stride + slot);
    uint64_t flag = w >> 62;
    bool cur = ((uint32_t)(w >> 32) & kEpochMask) == (epoch & kEpochMask);
    if (cur && flag == kFlagInclusive) return (uint32_t)w;
    if (cur && flag == kFlagAggregate) {
      excl = (uint32_t)w;
      --t;
    }
  }
  while (t >= 0) {
    uint64_t w[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
      w[j] = t - j >= 0 ? status_load(status + (uint64_t)(t - j) * stride + slot) : 0;
    int taken = 0;
    bool done = false;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      if (done || taken != j || t - j < 0) continue;
      uint64_t flag = w[j] >> 62;
      if (((uint32_t)(w[j] >> 32) & kEpochMask) != (epoch & kEpochMask) || flag == 0) continue;
      excl += (uint32_t)w[j];
      ++taken;
      done = flag == kFlagInclusive;
    }
    if (done) break;
    if (taken == 0) __builtin_amdgcn_s_sleep(1);
    t -= taken;
  }
  return excl;
}

// ------------------------------------------------------- wave primitives ---
__device__ inline uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// popcount(mask & lanes below this one)
__device__ inline uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ inline uint32_t wave_inclusive_sum(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(v, off, kWave);
    if (lane >= (uint32_t)off) v += y;
  }
  return v;
}

// Look-back over a single-slot status array by one whole wave: lane l reads
// tile (base - l), so one round trip inspects 64 predecessors.  Every lane
// returns the same exclusive prefix.
__device__ inline uint32_t wave_lookback(const uint64_t *status, uint32_t tile, uint32_t epoch) {
  const uint32_t lane = lane_id();
  uint32_t excl = 0;
  int64_t base = (int64_t)tile - 1;
  while (base >= 0) {
    const int64_t t = base - (int64_t)lane;
    bool ready = true, incl = true;
    uint32_t val = 0;
    if (t >= 0) {
      uint64_t w = status_load(status + t);
      uint64_t flag = w >> 62;
      ready = ((uint32_t)(w >> 32) & kEpochMask) == (epoch & kEpochMask) && flag != 0;
      incl = ready && flag == kFlagInclusive;
      val = (uint32_t)w;
    }
    const uint64_t incl_mask = __ballot(incl);
    const uint64_t not_ready = __ballot(!ready);
    const uint32_t stop = incl_mask ? (uint32_t)__builtin_ctzll(incl_mask) : 63u;
    const uint64_t needed = stop == 63u ? ~0ull : ((2ull << stop) - 1ull);
    if (not_ready & needed) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t part = lane <= stop ? val : 0u;
    part = wave_inclusive_sum(part);
    excl += __shfl(part, 63, kWave);
    if (incl_mask) break;
    base -= kWave;
  }
  return excl;
}

// Block-wide exclusive sum for a 256-thread block; `scratch` holds 4 words.
// Returns the exclusive prefix; *total receives the block sum.
__device__ inline uint32_t block256_exclusive_sum(uint32_t v, uint32_t *scratch,
                                                  uint32_t *total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t incl = wave_inclusive_sum(v);
  if (lane == 63) scratch[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    uint32_t s = scratch[w];
    before += (uint32_t)w < wave ? s : 0;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// Exclusive sum over a block of BT threads (BT/64 waves); scratch holds BT/64
// words.  Every thread of the block must call it.
template <int BT>
__device__ inline uint32_t block_exclusive_sum_bt(uint32_t v, uint32_t *scratch, uint32_t *total) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t incl = wave_inclusive_sum(v);
  if (lane == 63) scratch[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < BT / kWave; ++w) {
    const uint32_t x = scratch[w];
    before += (uint32_t)w < wave ? x : 0;
    all += x;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__device__ inline uint64_t wave_inclusive_sum64(uint64_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint64_t y = __shfl_up(v, off, kWave);
    if (lane >= (uint32_t)off) v += y;
  }
  return v;
}

// Exclusive sum over a block of W waves; `scratch` holds W words.
template <uint32_t W>
__device__ inline uint64_t block_exclusive_sum_waves(uint64_t v, uint64_t *scratch, uint64_t *total) {
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint64_t incl = wave_inclusive_sum64(v);
  if (lane == kWave - 1) scratch[wave] = incl;
  __syncthreads();
  uint64_t before = 0, all = 0;
#pragma unroll
  for (uint32_t w = 0; w < W; ++w) {
    const uint64_t x = scratch[w];
    before += w < wave ? x : 0;
    all += x;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// 64-bit variant for a 256-thread block; `scratch` holds 4 words.
__device__ inline uint64_t block_exclusive_sum64_lds(uint64_t v, uint64_t *scratch,
                                                     uint64_t *total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t incl = wave_inclusive_sum64(v);
  if (lane == 63) scratch[wave] = incl;
  __syncthreads();
  uint64_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    uint64_t s = scratch[w];
    before += (uint32_t)w < wave ? s : 0;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// ---------------------------------------------------------------- context ---
struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// Grows `b` to at least `bytes` (contents are not preserved).  While a
// stream capture is being recorded on this thread, growing fails instead
// (allocations are not stream work; workspaces are sized before capture).
int ensure(DeviceBuffer &b, size_t bytes);
void set_alloc_forbidden(bool forbidden);
void release(DeviceBuffer &b);

// Per-call device staging for host-memory arguments.
struct Staging {
  DeviceBuffer b[6];
  ~Staging() {
    for (auto &x : b) release(x);
  }
};

// Makes `dev` current for the lifetime of the guard.
struct DeviceGuard {
  int prev = -1;
  bool ok = true;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Entry-point prologue: rejects a null context and makes its device current.
#define LNT_CTX(ctx)                                                       \
  if (!(ctx)) return ::lnt::fail(LNT_ERR_INVALID_ARGUMENT, "null context"); \
  ::lnt::DeviceGuard _guard((ctx)->device);                                \
  if (!_guard.ok) return ::lnt::fail(LNT_ERR_DEVICE, "hipSetDevice failed")

inline int check_mem(int mem) {
  if (mem != LNT_MEM_HOST && mem != LNT_MEM_DEVICE)
    return fail(LNT_ERR_INVALID_ARGUMENT, "mem must be LNT_MEM_HOST or LNT_MEM_DEVICE");
  return LNT_OK;
}

struct SortTuning {
  int items_per_thread = 12;  // onesweep tile = sort_block * items
  int hist_blocks = 4096;
  int merge_items = 0;    // merge tile = merge_block * items; 0: 16 with 32-bit LDS keys, else 12
  int merge_block = 512;  // merge threads per workgroup (256 or 512)
  int spgemm_small = 2048;  // largest output column handled in LDS
  int narrow_keys = 1;  // 32-bit keys between passes when they fit
  int row_ranks = 1;    // lattice product: long columns sort row ranks when that saves passes
  int spgemm_pieces = 1;  // lattice product: long columns in row pieces (0: all through the global sort)
  int spgemm_single_copy = 1;  // lattice product: single-entry columns by the XCD-local copy kernel
  int digit_bits = 8;  // largest radix digit per onesweep pass (4..8)
  int sort_block = 1024;  // onesweep threads per block (256, 512, 1024): tile = sort_block * items
};

// Kernel kinds recorded by the profiler (names in api.hip).
enum ProfKind { kProfHistogram = 0, kProfOnesweep, kProfPartition, kProfMerge, kProfRelin,
                kProfLatticeMult, kProfKinds };

struct ProfRecord {
  int kind;
  hipEvent_t start, stop;
  double bytes;  // algorithmic bytes moved by this launch
};

}  // namespace lnt

struct lnt_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  int validate = 1;
  int capturing = 0;  // launches are being recorded into a graph: no host syncs
  uint32_t epoch = 0;
  lnt::SortTuning tuning;
  lnt::DeviceBuffer alt_keys, alt_vals, alt2_keys, alt2_vals;
  lnt::DeviceBuffer status;     // look-back words
  lnt::DeviceBuffer small;      // histograms, counters, flags
  lnt::DeviceBuffer aux;        // merge-path splits and other per-call scratch
  lnt::DeviceBuffer hist_part;  // per-block digit histograms
  lnt::DeviceBuffer relin_dev;  // relinearisation table of the current sort's radix passes
  lnt::DeviceBuffer spg_ws, spg_tmp, spg_big, spg_rank, spg_cp, spg_plan, spg_piece, spg_chunk;  // lattice-product workspaces
  lnt::DeviceBuffer host_tmp_a, host_tmp_b, host_tmp_c;  // staging for host-memory calls
  lnt::DeviceBuffer exec_buf[6];  // operand copies of a whole contraction (keys, values of both), timing output
  void *model = nullptr;          // lnt_model used by LNT_STRAT_AUTO (not owned)
  uint32_t *pinned = nullptr;   // small pinned host buffer for counts/flags
  int profiling = 0;
  std::vector<lnt::ProfRecord> prof;   // recorded launches (events owned)
  std::vector<hipEvent_t> event_pool;  // recycled events
  hipEvent_t check_event = nullptr;    // marks the sort's key-range flag copy
  void *noalloc_scratch = nullptr;     // this device's static merge scratch (lnt_merge_into_noalloc)
};

namespace lnt {

// Layout of ctx->small (uint32 words).
constexpr size_t kSmallHist = 0;                  // [8][256] digit histograms
constexpr size_t kSmallCounters = 8 * 256;        // [16] tile counters
constexpr size_t kSmallFlags = kSmallCounters + 16;  // [16] error flags / totals
constexpr size_t kSmallWords = kSmallFlags + 16;
constexpr int kMaxHistBlocks = 4096;

// Profiling: prof_begin records a start event (returns a slot, or -1 when
// profiling is off); prof_end records the stop event of that slot.
int prof_begin(lnt_context ctx, int kind, double bytes);
void prof_end(lnt_context ctx, int slot);
void prof_set_bytes(lnt_context ctx, int slot, double bytes);

// Next epoch for a look-back pass; clears the status array on wrap-around.
int next_epoch(lnt_context ctx, uint32_t *epoch);
int ensure_status(lnt_context ctx, size_t words);

// Kernel launchers (sort.hip, merge.hip).  All pointers are device pointers.
int launch_fill_u32(lnt_context ctx, uint32_t *p, uint32_t value, size_t n);
// sorted_low: number of low key bits (after the relinearisation) whose order
// the input already carries -- see sorted_low_bits; the radix passes then
// cover only the bits above them.
int sort_device(lnt_context ctx, uint64_t *keys, double *vals, uint32_t n,
                uint64_t bound, const Relin *relin, uint32_t sorted_low = 0);

// For keys sorted in cur_seq, relinearised into new_seq: let the first J
// modes of new_seq (the fastest) appear in cur_seq in the same relative
// order, and let their dims multiply to P = 2^b * odd.  Two relinearised keys
// equal above bit b then agree on every other mode, so the old key orders
// them by those J modes in the new significance, i.e. by the low b bits: the
// input order is already right there and a stable LSD sort of the bits from
// b up gives the full order.  Returns b for the largest such J.
uint32_t sorted_low_bits(int order, const int64_t *dims, const int32_t *cur_seq, const int32_t *new_seq);

// Equality of two sorted operands in one linearisation (compare.hip).
int compare_device(lnt_context ctx, const uint64_t *ak, const double *av, uint32_t na, const uint64_t *bk,
                   const double *bv, uint32_t nb, int mode, double prec, int *equal);
int relin_device(lnt_context ctx, uint64_t *keys, uint32_t n, const Relin &relin);
int relin_device_large(lnt_context ctx, uint64_t *keys, uint64_t n, const Relin &relin);
int sort_device_large(lnt_context ctx, uint64_t *keys, double *vals, uint64_t n, uint64_t bound,
                      const Relin *relin);
int lattice_mult_device(lnt_context ctx, const uint64_t *ak, const double *av, uint64_t na,
                        const int64_t a_dims[3], const uint64_t *bk, const double *bv,
                        uint64_t nb, const int64_t b_dims[3], int strategy, uint64_t *ck,
                        double *cv, uint64_t capacity, uint64_t *nc_host, uint64_t *nprod_host,
                        uint64_t **ck_alloc, double **cv_alloc);
int lattice_nprod_device(lnt_context ctx, const uint64_t *ak, uint64_t na, const int64_t a_dims[3],
                         const uint64_t *bk, uint64_t nb, const int64_t b_dims[3],
                         uint64_t *nprod_host);
int merge_device(lnt_context ctx, const uint64_t *ak, const double *av,
                 uint32_t na, const uint64_t *bk, const double *bv, uint32_t nb,
                 int op, uint64_t *ck, double *cv, uint64_t capacity,
                 uint32_t *nc_host);
// The static scratch of the no-allocation merge (merge.hip): its address on
// the context's device, looked up once per device.
int noalloc_scratch_init(lnt_context ctx);
int merge_noalloc_device(lnt_context ctx, const uint64_t *ak, const double *av, uint32_t na, const uint64_t *bk,
                         const double *bv, uint32_t nb, int op, uint64_t *ck, double *cv, uint64_t capacity,
                         uint32_t *nc_host);
int merge_device_large(lnt_context ctx, const uint64_t *ak, const double *av, uint64_t na, const uint64_t *bk,
                       const double *bv, uint64_t nb, int op, uint64_t *ck, double *cv, uint64_t capacity,
                       uint64_t *nc_host);
int sort_merge_device(lnt_context ctx, const uint64_t *ak, const double *av, uint32_t na, const uint64_t *bk,
                      const double *bv, uint32_t nb, int op, uint64_t bound, uint64_t *ck, double *cv,
                      uint64_t capacity, uint64_t *nc_host);

}  // namespace lnt
