// batch.hip -- many independent sparse contractions sharing every launch,
// for gfx950.
//
// Every instance is the lattice product of spgemm.hip, C(:,j,t) =
// sum_k B(k,j,t) * A(:,k,t) (SparseLatticeBase.h:1854-1990), after both operands
// are re-sorted into lattice order (SparseMIABase.h:590-645 sort(seq)).  The
// reference runs one expression at a time; here one pipeline serves a whole
// batch and never waits for the host between its first and last launch, so it
// can be recorded once into a hipGraph and replayed.
//
//   1. relinearise: every key into its instance's lattice order plus the
//      instance's base, so that one radix sort of all instances at once leaves
//      each instance's entries, sorted, at the offsets they came in at;
//   2. columns: per B entry the A column range (binary search inside its
//      instance), the output column and segment heads; scans give product
//      offsets and segment ids (segment = one output column of one instance);
//   3. classify: single-entry segments (a scaled copy of one A column), small
//      segments (<= small_max products: grouped into work items of about
//      item_products products, sorted in LDS together) and large segments
//      (processed alone, in row windows of about `window` products);
//   4. symbolic pass: unique rows per segment, then a scan gives every segment
//      its output offset;
//   5. numeric pass: the same work again, writing C straight to its offsets
//      (no intermediate buffer: HBM holds only the inputs, the column tables
//      and C).
// Steps 4 and 5 run on a persistent grid that takes work items from atomic
// tickets (a work queue).  Equal rows are summed in ascending B-row order --
// the order of the sort strategy of spgemm.hip -- except for one row met by
// more than kCap products of one column, which is summed by a tree.
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lnt_batch.h"
#include "../../include/lnt_dispatch.h"
#include "lnt_internal.h"

namespace lnt {

template <typename T>
int exclusive_scan(lnt_context ctx, const T *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                   DeviceBuffer &tmp, size_t tmp_offset);
size_t scan_workspace_words(uint64_t n);

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kCap = 2048;      // LDS elements of one sorted group
constexpr int kPer = kCap / kBlock;  // sorted elements per thread
constexpr uint32_t kMaxSegsPerItem = 1024 + 1;
constexpr int kRowBits = 40;         // small items sort (segment, row) in one word
constexpr uint64_t kRowMask = (uint64_t(1) << kRowBits) - 1;
// Dense segments (batch_dense_kernel): the tallest lattice the dense
// accumulator takes, the height up to which its smaller variant serves, and
// entries per chunk.
constexpr uint32_t kDenseH = 8192;
constexpr uint32_t kDenseSmallH = 2048;
constexpr uint32_t kDenseEnt = 64;

// Counters and totals (uint64 words of the batch's control block).
enum {
  kCtrNprod = 0,    // total products (scan total)
  kCtrNseg,         // segments
  kCtrNsmall,       // small segments
  kCtrSmallProd,    // products of small segments
  kCtrNc,           // nnz of all C
  kCtrN3,           // large segments
  kCtrTicketLarge,
  kCtrTicketSmall,
  kCtrTicketDense,  // symbolic pass, then the numeric pass of the smaller heights
  kCtrTicketDense2, // numeric pass of the taller ones
  kCtrKeyErr,       // an input key outside its instance's dimensionality
  kCtrOverflow,     // C larger than the output capacity
  kCtrN3w,          // large segments taken by the windows
  kCtrPieces,       // piece slots the batch needs
  kCtrTicketPiece,  // pieces written straight into C take tickets (look-back order)
  kCtrNdense,       // dense segments: (lattice height <= kDenseSmallH) | (taller) << 32
  kCtrWords = 16
};

struct InstDev {
  uint64_t a_off, na, b_off, nb;  // packed entry ranges
  uint64_t a_base, b_base;        // composite key bases
  uint64_t a_dim, b_dim;          // key bounds
  uint64_t h, w, bw;              // A h x w x d, B w x bw x d
  Divider div_w, div_wbw;         // B lattice key -> (k, j, t)
};

__device__ inline uint64_t lower_bound_u64(const uint64_t *a, uint64_t lo, uint64_t hi, uint64_t x) {
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Largest i in [lo, hi) with off[i] <= x (off ascending, off[lo] <= x).
template <typename T>
__device__ inline uint64_t upper_index(const T *off, uint64_t lo, uint64_t hi, uint64_t x) {
  while (hi - lo > 1) {
    uint64_t mid = (lo + hi) >> 1;
    if ((uint64_t)off[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t *scratch) {
  uint64_t total;
  block_exclusive_sum64_lds(v, scratch, &total);
  return total;
}

__device__ inline uint64_t block_min_u64(uint64_t v, uint64_t *scratch) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t y = __shfl_xor(v, off, kWave);
    v = y < v ? y : v;
  }
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  uint64_t m = scratch[0];
#pragma unroll
  for (int w = 1; w < kBlock / kWave; ++w) m = scratch[w] < m ? scratch[w] : m;
  __syncthreads();
  return m;
}

__device__ inline uint32_t bit_width(uint64_t x) { return x ? 64u - (uint32_t)__clzll(x) : 0u; }

__device__ inline uint64_t block_max_u64(uint64_t v, uint64_t *scratch) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t y = __shfl_xor(v, off, kWave);
    v = y > v ? y : v;
  }
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  uint64_t m = scratch[0];
#pragma unroll
  for (int w = 1; w < kBlock / kWave; ++w) m = scratch[w] > m ? scratch[w] : m;
  __syncthreads();
  return m;
}

// Tree sum of one double per thread (deterministic for a fixed assignment).
__device__ inline double block_sum_f64(double v, double *scratch) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double s = (scratch[0] + scratch[1]) + (scratch[2] + scratch[3]);
  __syncthreads();
  return s;
}

// Bitonic sort of s_key[0, P) with s_pos as tie-break; P a power of two in
// [2, kCap].  Each step, every thread handles P/2/kBlock compare-exchanges.
__device__ inline void bitonic_key_pos(uint64_t *s_key, uint16_t *s_pos, uint32_t P) {
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t q = threadIdx.x; q < (P >> 1); q += kBlock) {
        const uint32_t lo = ((q & ~(j - 1)) << 1) | (q & (j - 1));
        const uint32_t hi = lo + j;
        const uint64_t kl = s_key[lo], kh = s_key[hi];
        const uint16_t pl = s_pos[lo], ph = s_pos[hi];
        const bool gt = kl > kh || (kl == kh && pl > ph);
        const bool up = (lo & k) == 0;
        if (gt == up) {
          s_key[lo] = kh;
          s_key[hi] = kl;
          s_pos[lo] = ph;
          s_pos[hi] = pl;
        }
      }
      __syncthreads();
    }
  }
}

// Stable LSD radix sort (8-bit digits) of (key, pos) pairs in LDS by bits
// [0, lo_bits) and [hi_shift, hi_shift + hi_bits) of the key (all other bits
// zero); n <= kCap; the result is left in (k0, p0) when the pass count is
// even, else in (k1, p1) (returned: 0 or 1).  Element arrangement:
// wave w owns elements [w*512, w*512+512), lane l of slot u holds
// w*512 + u*64 + l, so ranking slot by slot inside each wave and waves in
// order keeps equal keys in element order.  Peers with the same digit are
// found by ballot matching; per-wave digit counters live in s_cnt[4][256].
template <typename K>
__device__ inline int lds_radix_sort(K *k0, uint16_t *p0, K *k1, uint16_t *p1, uint32_t n,
                                     uint32_t lo_bits, uint32_t hi_shift, uint32_t hi_bits, uint32_t *s_cnt,
                                     uint32_t *s_scan) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t first_idx = (int32_t)(wave * 512);
  const int slots = n > (uint32_t)first_idx ? (int)min(8u, (n - first_idx + 63) / 64) : 0;
  int buf = 0;
  const uint32_t lo_passes = (lo_bits + 7) / 8, passes = lo_passes + (hi_bits + 7) / 8;
  for (uint32_t pass = 0; pass < passes; ++pass) {
    const uint32_t shift = pass < lo_passes ? 8 * pass : hi_shift + 8 * (pass - lo_passes);
    K *kin = buf ? k1 : k0, *kout = buf ? k0 : k1;
    uint16_t *pin = buf ? p1 : p0, *pout = buf ? p0 : p1;
    for (uint32_t i = tid; i < 4 * 256; i += kBlock) s_cnt[i] = 0;
    __syncthreads();
    K key[8];
    uint16_t pos[8];
    uint32_t rank[8], dig[8];
    const uint64_t lt = (1ull << lane) - 1;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (u >= slots) break;
      const uint32_t idx = first_idx + u * 64 + lane;
      key[u] = idx < n ? kin[idx] : (K)~(K)0;
      pos[u] = idx < n ? pin[idx] : 0;
      const uint32_t d = (uint32_t)(key[u] >> shift) & 255u;
      // Lanes that differ from this one in some digit bit: per bit, m = the
      // bit sign-extended, and x |= ballot ^ m (v_bitop3 truth table 0xF6 =
      // src0 | (src1 ^ src2)), one instruction per half of the wave.
      uint32_t x_lo = 0, x_hi = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)d, b, 1);  // 0 or ~0
        const uint64_t bal = __ballot(m != 0);
        x_lo = __builtin_amdgcn_bitop3_b32(x_lo, (uint32_t)bal, m, 0xF6);
        x_hi = __builtin_amdgcn_bitop3_b32(x_hi, (uint32_t)(bal >> 32), m, 0xF6);
      }
      const uint64_t match = ~(((uint64_t)x_hi << 32) | x_lo);
      const uint32_t r = (uint32_t)__popcll(match & lt);
      // Every lane of a digit group reads the count before any lane writes
      // it, and all write the same new count: no leader, no branch.
      const uint32_t base = s_cnt[wave * 256 + d];
      rank[u] = base + r;
      dig[u] = d;
      s_cnt[wave * 256 + d] = base + (uint32_t)__popcll(match);
    }
    __syncthreads();
    {
      const uint32_t d = tid;  // 256 threads <-> 256 digits
      const uint32_t c0 = s_cnt[d], c1 = s_cnt[256 + d], c2 = s_cnt[512 + d], c3 = s_cnt[768 + d];
      uint32_t tot;
      const uint32_t base = block256_exclusive_sum(c0 + c1 + c2 + c3, s_scan, &tot);
      s_cnt[d] = base;
      s_cnt[256 + d] = base + c0;
      s_cnt[512 + d] = base + c0 + c1;
      s_cnt[768 + d] = base + c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (u >= slots) break;
      const uint32_t dst = s_cnt[wave * 256 + dig[u]] + rank[u];
      kout[dst] = key[u];
      pout[dst] = pos[u];
    }
    __syncthreads();
    buf ^= 1;
  }
  return buf;
}

// Instance of element p, given that this block's elements lie in instances
// [lo, hi] (found once per block).
__device__ inline uint32_t instance_of(const uint64_t *off, uint32_t lo, uint32_t hi, uint64_t p) {
  return lo == hi ? lo : (uint32_t)upper_index(off, lo, (uint64_t)hi + 1, p);
}

// ------------------------------------------------------------------ 1. keys
// Composite key = instance base + key relinearised into lattice order; values
// copied alongside (the caller's operands stay untouched).
__global__ __launch_bounds__(kBlock) void batch_relin_kernel(
    const int64_t *__restrict__ keys, const double *__restrict__ vals, uint64_t n,
    const uint64_t *__restrict__ off, uint32_t n_inst, const InstDev *__restrict__ inst, int which,
    const RelinDev *__restrict__ relin, uint64_t *__restrict__ out_keys, double *__restrict__ out_vals,
    uint64_t *__restrict__ ctr) {
  __shared__ uint32_t s_lo, s_hi;
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < n; base += (uint64_t)gridDim.x * kBlock) {
    if (threadIdx.x == 0) {
      uint64_t last = (base + kBlock < n ? base + kBlock : n) - 1;
      s_lo = (uint32_t)upper_index(off, 0, n_inst, base);
      s_hi = (uint32_t)upper_index(off, 0, n_inst, last);
    }
    __syncthreads();
    const uint32_t lo = s_lo, hi = s_hi;
    __syncthreads();
    const uint64_t p = base + threadIdx.x;
    if (p >= n) continue;
    const uint32_t i = instance_of(off, lo, hi, p);
    const int64_t raw = keys[p];
    const uint64_t bound = which ? inst[i].b_dim : inst[i].a_dim;
    uint64_t k[1] = {raw < 0 || (uint64_t)raw >= bound ? 0 : (uint64_t)raw};
    if (raw < 0 || (uint64_t)raw >= bound) atomicOr((unsigned long long *)&ctr[kCtrKeyErr], 1ull);
    relin_keys<1>(relin + i, k);
    out_keys[p] = (which ? inst[i].b_base : inst[i].a_base) + k[0];
    if (vals) out_vals[p] = vals[p];
  }
}

// --------------------------------------------------------------- 2. columns
__global__ __launch_bounds__(kBlock) void batch_entry_kernel(
    const uint64_t *__restrict__ ak, const uint64_t *__restrict__ bk, uint64_t nb,
    const uint64_t *__restrict__ b_off, uint32_t n_inst, const InstDev *__restrict__ inst,
    uint64_t *__restrict__ a_lo, uint32_t *__restrict__ a_len, uint64_t *__restrict__ a_base,
    uint64_t *__restrict__ colkey, uint32_t *__restrict__ einst, uint32_t *__restrict__ head) {
  __shared__ uint32_t s_lo, s_hi;
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < nb; base += (uint64_t)gridDim.x * kBlock) {
    if (threadIdx.x == 0) {
      uint64_t last = (base + kBlock < nb ? base + kBlock : nb) - 1;
      s_lo = (uint32_t)upper_index(b_off, 0, n_inst, base);
      s_hi = (uint32_t)upper_index(b_off, 0, n_inst, last);
    }
    __syncthreads();
    const uint32_t lo = s_lo, hi = s_hi;
    __syncthreads();
    const uint64_t p = base + threadIdx.x;
    if (p >= nb) continue;
    const uint32_t i = instance_of(b_off, lo, hi, p);
    const InstDev &I = inst[i];
    const uint64_t key = bk[p] - I.b_base;
    const uint64_t t = divide(key, I.div_wbw);
    const uint64_t rem = key - t * I.div_wbw.d;
    const uint64_t j = divide(rem, I.div_w);
    const uint64_t k = rem - j * I.w;
    const uint64_t start = I.a_base + (t * I.w + k) * I.h;
    const uint64_t a_end = I.a_off + I.na;
    const uint64_t l = lower_bound_u64(ak, I.a_off, a_end, start);
    const uint64_t r = lower_bound_u64(ak, l, a_end, start + I.h);
    const uint64_t col = t * I.bw + j;
    a_lo[p] = l;
    a_len[p] = (uint32_t)(r - l);
    a_base[p] = start;
    colkey[p] = I.h * col;
    einst[p] = i;
    uint32_t hd = 1;
    if (p > I.b_off) {
      const uint64_t pk = bk[p - 1] - I.b_base;
      const uint64_t pt = divide(pk, I.div_wbw);
      const uint64_t pj = divide(pk - pt * I.div_wbw.d, I.div_w);
      hd = (pt * I.bw + pj) != col;
    }
    head[p] = hd;
  }
}

__global__ void batch_seg_first_kernel(const uint32_t *__restrict__ head, const uint64_t *__restrict__ seg_idx,
                                       uint64_t nb, uint64_t *__restrict__ seg_first) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < nb; i += (uint64_t)gridDim.x * blockDim.x)
    if (head[i]) seg_first[seg_idx[i]] = i;
}

__device__ inline uint64_t seg_last(const uint64_t *seg_first, uint64_t s, uint64_t nseg, uint64_t nb) {
  return s + 1 < nseg ? seg_first[s + 1] : nb;
}

// -------------------------------------------------------------- 3. classify
// kind: 0 no products, 1 single B entry, 2 small, 3 large, 4 and 5 dense (an
// instance with the dense accumulator, lattice height up to kDenseSmallH or
// taller: see batch_dense_kernel; dense_flag packs the two kinds as 1 and
// 1 << 32, so that one scan ranks both).  Arrays are written for all nb slots
// (zero past the last segment) so that scans over nb elements see only real
// segments.
__global__ void batch_classify_kernel(const uint64_t *__restrict__ seg_first, const uint64_t *__restrict__ prod_off,
                                      const uint32_t *__restrict__ einst, const InstDev *__restrict__ inst,
                                      const uint32_t *__restrict__ inst_dense, uint64_t nb, uint32_t small_max,
                                      uint32_t dense_min, uint64_t *__restrict__ ctr,
                                      uint32_t *__restrict__ kind, uint32_t *__restrict__ is_small,
                                      uint64_t *__restrict__ seg_u, uint32_t *__restrict__ is_large,
                                      uint64_t *__restrict__ dense_flag) {
  const uint64_t nseg = ctr[kCtrNseg];
  uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; s < nb; s += (uint64_t)gridDim.x * blockDim.x) {
    if (s >= nseg) {
      kind[s] = 0;
      is_small[s] = 0;
      is_large[s] = 0;
      dense_flag[s] = 0;
      seg_u[s] = 0;
      continue;
    }
    const uint64_t first = seg_first[s], last = seg_last(seg_first, s, nseg, nb);
    const uint64_t np = prod_off[last] - prod_off[first];
    const uint32_t i = einst[first];
    const uint64_t h = inst[i].h;
    uint32_t k;
    if (np == 0) k = 0;
    else if (last - first == 1) k = 1;
    else if (inst_dense[i] && h <= kDenseH && np >= dense_min && np >= 8 * (last - first))
      k = h <= kDenseSmallH ? 4 : 5;
    else if (np <= small_max && h <= kRowMask) k = 2;
    else k = 3;
    kind[s] = k;
    is_small[s] = k == 2;
    is_large[s] = k == 3;
    dense_flag[s] = k == 4 ? 1ull : k == 5 ? (1ull << 32) : 0ull;
    seg_u[s] = k == 1 ? np : 0;
  }
}

// What a dense segment's workgroup needs first, in one load: its entries
// [first, first + nent), products [p0, p1), segment id and lattice height.
struct DenseDesc {
  uint64_t first, p0, p1;
  uint32_t seg, nent, h, pad;
};

// The descriptors of the dense segments: the smaller heights from the front
// of `desc` (nb + 1 slots), the taller ones from the back, each in ascending
// segment order.
__global__ void batch_dense_list_kernel(const uint64_t *__restrict__ dense_flag, const uint64_t *__restrict__ rank,
                                        const uint64_t *__restrict__ seg_first, const uint64_t *__restrict__ prod_off,
                                        const uint32_t *__restrict__ einst, const InstDev *__restrict__ inst,
                                        const uint64_t *__restrict__ ctr, uint64_t nb,
                                        DenseDesc *__restrict__ desc) {
  const uint64_t nseg = ctr[kCtrNseg];
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t f = dense_flag[s];
    if (!f) continue;
    const uint64_t r = rank[s];
    const uint64_t idx = (f & 1) ? (r & 0xFFFFFFFFull) : nb - (r >> 32);
    const uint64_t first = seg_first[s], last = seg_last(seg_first, s, nseg, nb);
    DenseDesc d;
    d.first = first;
    d.p0 = prod_off[first];
    d.p1 = prod_off[last];
    d.seg = (uint32_t)s;
    d.nent = (uint32_t)(last - first);
    d.h = (uint32_t)inst[einst[first]].h;
    d.pad = 0;
    desc[idx] = d;
  }
}

// The large segments in ascending order (their pieces then come in output
// order).
__global__ void batch_large_list_kernel(const uint32_t *__restrict__ is_large, const uint64_t *__restrict__ rank,
                                        uint64_t nb, uint32_t *__restrict__ list3) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nb; s += (uint64_t)gridDim.x * blockDim.x)
    if (is_large[s]) list3[rank[s]] = (uint32_t)s;
}

// Ordered list of small segments and their product counts (np_list must be
// zero beyond the last small segment).
__global__ void batch_small_list_kernel(const uint32_t *__restrict__ is_small, const uint64_t *__restrict__ small_rank,
                                        const uint64_t *__restrict__ seg_first, const uint64_t *__restrict__ prod_off,
                                        uint64_t nb, const uint64_t *__restrict__ ctr,
                                        uint32_t *__restrict__ small_list, uint32_t *__restrict__ np_list) {
  const uint64_t nseg = ctr[kCtrNseg];
  uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; s < nseg; s += (uint64_t)gridDim.x * blockDim.x) {
    if (!is_small[s]) continue;
    const uint64_t r = small_rank[s];
    const uint64_t first = seg_first[s], last = seg_last(seg_first, s, nseg, nb);
    small_list[r] = (uint32_t)s;
    np_list[r] = (uint32_t)(prod_off[last] - prod_off[first]);
  }
}

// Product p (0-based inside the segment [first, last)) -> B entry e and A index.
__device__ inline void locate(const uint64_t *prod_off, uint64_t first, uint64_t last, uint64_t p,
                              const uint64_t *a_lo, uint64_t *e_out, uint64_t *a_idx) {
  const uint64_t base = prod_off[first];
  const uint64_t e = upper_index(prod_off, first, last, base + p);
  *e_out = e;
  *a_idx = a_lo[e] + (base + p - prod_off[e]);
}

// Work tickets taken K at a time by one thread of a workgroup: every atomic
// on one address is serialised device-wide (a few ns each), so a counter
// bumped once per small work item caps the rate at which workgroups can
// start items; K consecutive items keep the workgroups on neighbouring items
// (their operands share cache lines) and cut the atomics K-fold.  The
// workgroup runs its items in order; none waits on another item.
template <uint32_t K>
struct TicketBlock {
  uint64_t cur = 0, end = 0;
  __device__ uint64_t next(unsigned long long *counter) {
    if (cur == end) {
      cur = atomicAdd(counter, (unsigned long long)K);
      end = cur + K;
    }
    return cur++;
  }
};

struct ColumnArgs {
  const uint64_t *ak;
  const double *av, *bv;
  const uint64_t *a_lo, *a_base, *colkey, *prod_off, *seg_first, *out_off;
  const uint32_t *a_len, *einst;
  const InstDev *inst;
  uint64_t nb;
  uint64_t *seg_u;  // symbolic output
  uint64_t *ck;     // numeric output
  double *cv;
  uint64_t capacity;
};

// LDS hash set of 64-bit keys: the symbolic pass counts distinct keys
// without sorting (at most kCap keys in a table of 2*kCap slots).
constexpr uint32_t kHashSlots = 2 * kCap;
constexpr unsigned long long kEmptyKey = ~0ull;

// The same set of 32-bit keys (every key below 0xFFFFFFFF), in half the LDS.
__device__ inline bool hash_insert(unsigned int *table, uint32_t key) {
  uint32_t slot = (uint32_t)(((uint64_t)key * 0x9E3779B97F4A7C15ull) >> 52) & (kHashSlots - 1);
  for (;;) {
    const unsigned int prev = atomicCAS(&table[slot], 0xFFFFFFFFu, key);
    if (prev == 0xFFFFFFFFu) return true;
    if (prev == key) return false;
    slot = (slot + 1) & (kHashSlots - 1);
  }
}

__device__ inline bool hash_insert(unsigned long long *table, uint64_t key) {
  uint32_t slot = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 52) & (kHashSlots - 1);
  for (;;) {
    const unsigned long long prev = atomicCAS(&table[slot], kEmptyKey, (unsigned long long)key);
    if (prev == kEmptyKey) return true;
    if (prev == key) return false;
    slot = (slot + 1) & (kHashSlots - 1);
  }
}

// Position inside one B entry's run of products: entry e, the A index of
// product g of the flat product numbering is alo + g.
struct EntryCursor {
  uint64_t e, end, alo, abase;
  double b;
};

template <bool NUMERIC>
__device__ inline void load_entry(const ColumnArgs &A, uint64_t e, EntryCursor &c) {
  c.e = e;
  c.end = A.prod_off[e + 1];
  c.alo = A.a_lo[e] - A.prod_off[e];
  c.abase = A.a_base[e];
  c.b = NUMERIC ? A.bv[e] : 0.0;
}

// First x in [lo, hi) with a[x] >= key (hi if none), searched by the G lanes
// of a group (G a power of two, 2..32) in rounds of G probes: a[x] < key
// holds below the answer and fails from it on, so one ballot over G evenly
// spaced x narrows the range (G+1)-fold.  Every lane of the wave calls it
// with the same G; a group with `active` false returns lo.
__device__ inline uint64_t group_lower_bound(const uint64_t *__restrict__ a, uint64_t lo, uint64_t hi, uint64_t key,
                                             bool active, uint32_t G) {
  const uint32_t lane = threadIdx.x & (kWave - 1), gl = lane & (G - 1), shift = lane - gl;
  const uint64_t mask = (1ull << G) - 1;
  if (!active) hi = lo;
  while (__ballot(hi > lo)) {
    const uint64_t w = hi - lo;
    // Probes inside [lo, hi): lo + gl when the range fits the group, else
    // strictly increasing points.
    const uint64_t x = w <= G ? lo + gl : lo + ((uint64_t)(gl + 1) * w) / (G + 1);
    const bool live = w > G || gl < w;
    const bool below = live && a[x] < key;
    const uint32_t trues = (uint32_t)__popcll((__ballot(below) >> shift) & mask);
    const uint64_t x_last = __shfl(x, shift + (trues ? trues - 1 : 0), kWave);
    const uint64_t x_fail = __shfl(x, shift + (trues < G ? trues : G - 1), kWave);
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

// ------------------------------------------------ 4/5. small segments, items
// Work item t = the small segments whose compact product offset lies in
// [t*G, (t+1)*G): at most G segments and G + small_max - 1 products.  Each
// thread expands kPer consecutive products (one search for its first entry,
// then a walk).  Symbolic: distinct (segment, row) keys counted in an LDS
// hash set.  Numeric: the products are sorted together by (segment, row,
// position) and equal runs summed in position order.
template <bool NUMERIC, typename K>
__global__ __launch_bounds__(kBlock) void batch_small_kernel(ColumnArgs A, const uint32_t *__restrict__ small_list,
                                                             const uint64_t *__restrict__ sp_off, uint32_t G,
                                                             uint32_t rbits, uint64_t *__restrict__ ctr) {
  // Sort keys (segment of the item << rbits) | row in K words: 32 bits when
  // rbits plus the item's segment bits fit (the batch's rows are below
  // 2^rbits).  In
  // the numeric pass the per-segment rank bases share LDS with the sort's
  // digit counters (they are written after the sort), and the item's product
  // offsets are 16-bit (an item holds fewer than 2 * kCap products).
  using HK = typename std::conditional<sizeof(K) == 4, unsigned int, unsigned long long>::type;
  __shared__ HK s_hash[NUMERIC ? 1 : kHashSlots];
  __shared__ K s_key[NUMERIC ? kCap : 1];
  __shared__ uint16_t s_pos[NUMERIC ? kCap : 1];
  __shared__ K s_key2[NUMERIC ? kCap : 1];
  __shared__ uint16_t s_pos2[NUMERIC ? kCap : 1];
  __shared__ uint32_t s_rcnt[NUMERIC ? (4 * 256 > kMaxSegsPerItem ? 4 * 256 : kMaxSegsPerItem) : kMaxSegsPerItem];
  __shared__ uint64_t s_red[kBlock / kWave];
  __shared__ double s_val[NUMERIC ? kCap : 1];
  __shared__ uint32_t s_seg[kMaxSegsPerItem];
  __shared__ uint16_t s_spo[kMaxSegsPerItem + 1];
  uint32_t *const s_cnt = s_rcnt;  // symbolic: distinct rows; numeric: rank base
  __shared__ uint32_t s_scan[kBlock / kWave];
  __shared__ uint64_t s_item, s_rb, s_re;
  const uint32_t tid = threadIdx.x;
  const uint64_t nseg = ctr[kCtrNseg];
  const uint64_t n_small = ctr[kCtrNsmall];
  const uint64_t total = ctr[kCtrSmallProd];
  const uint64_t n_items = (total + G - 1) / G;
  TicketBlock<4> tickets;
  for (;;) {
    if (tid < kWave) {
      // The item's first and last small segments: two searches over the
      // segments' product offsets by 32 lanes each, in rounds of 32 probes.
      uint64_t item = 0;
      if (tid == 0) item = tickets.next((unsigned long long *)&ctr[kCtrTicketSmall]);
      item = __shfl(item, 0, kWave);
      const uint64_t x = group_lower_bound(sp_off, 0, n_small, (item + (tid >= 32 ? 1 : 0)) * G, item < n_items, 32);
      if (tid == 0) {
        s_item = item;
        s_rb = x;
      }
      if (tid == 32) s_re = x;
    }
    __syncthreads();
    const uint64_t item = s_item;
    if (item >= n_items) break;
    const uint64_t rb = s_rb, re = s_re;
    const uint32_t nr = (uint32_t)(re - rb);
    const uint64_t pbase = rb < n_small ? sp_off[rb] : total;
    for (uint32_t q = tid; q <= nr; q += kBlock) {
      if (q < nr) {
        s_seg[q] = small_list[rb + q];
        s_cnt[q] = 0;
      }
      s_spo[q] = (uint16_t)((rb + q < n_small ? sp_off[rb + q] : total) - pbase);
    }
    if (!NUMERIC)
      for (uint32_t i = tid; i < kHashSlots; i += kBlock) s_hash[i] = (HK)kEmptyKey;
    __syncthreads();
    const uint32_t np = s_spo[nr];
    const uint32_t p0 = tid * kPer;
    uint64_t maxrow = 0;
    if (p0 < np) {
      uint32_t q = (uint32_t)upper_index(s_spo, 0, nr, p0);
      uint64_t s = s_seg[q];
      uint64_t first = A.seg_first[s], last = seg_last(A.seg_first, s, nseg, A.nb);
      uint64_t sbase = A.prod_off[first];
      uint32_t seg_end = s_spo[q + 1];
      uint64_t g = sbase + (p0 - s_spo[q]);
      EntryCursor c;
      load_entry<NUMERIC>(A, upper_index(A.prod_off, first, last, g), c);
      for (int u = 0; u < kPer; ++u) {
        const uint32_t p = p0 + u;
        if (p >= np) break;
        if (p >= seg_end) {  // next small segment of the item (each has >= 1 product)
          ++q;
          s = s_seg[q];
          first = A.seg_first[s];
          last = seg_last(A.seg_first, s, nseg, A.nb);
          sbase = A.prod_off[first];
          seg_end = s_spo[q + 1];
          g = sbase + (p - s_spo[q]);
          load_entry<NUMERIC>(A, upper_index(A.prod_off, first, last, g), c);
        } else if (g >= c.end) {  // next entry with products
          load_entry<NUMERIC>(A, upper_index(A.prod_off, c.e + 1, last, g), c);
        }
        const uint64_t ai = c.alo + g;
        const uint64_t row = A.ak[ai] - c.abase;
        const uint64_t key = ((uint64_t)q << rbits) | row;
        maxrow = row > maxrow ? row : maxrow;
        if (NUMERIC) {
          s_key[p] = (K)key;
          s_pos[p] = (uint16_t)p;
          s_val[p] = c.b * A.av[ai];
        } else if (hash_insert(s_hash, (K)key)) {
          atomicAdd(&s_cnt[q], 1u);
        }
        ++g;
      }
    }
    if (!NUMERIC) {
      __syncthreads();
      for (uint32_t q = tid; q < nr; q += kBlock) A.seg_u[s_seg[q]] = s_cnt[q];
      __syncthreads();
      continue;
    }
    const uint32_t rowbits = bit_width(block_max_u64(maxrow, s_red));
    const int which = lds_radix_sort<K>(s_key, s_pos, s_key2, s_pos2, np, rowbits, rbits,
                                        nr > 1 ? bit_width(nr - 1) : 0, s_rcnt, s_scan);
    const K *kk = which ? s_key2 : s_key;
    const uint16_t *pp = which ? s_pos2 : s_pos;
    // Heads of equal-key runs, ranked; the first element of every segment
    // records its rank so that ranks can be made segment-relative.
    uint32_t flags = 0, cnt = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const uint32_t p = tid * kPer + u;
      const bool h = p < np && (p == 0 || kk[p] != kk[p - 1]);
      flags |= (h ? 1u : 0u) << u;
      cnt += h;
    }
    uint32_t heads;
    uint32_t rank = block256_exclusive_sum(cnt, s_scan, &heads);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const uint32_t p = tid * kPer + u;
      if ((flags >> u) & 1u) {
        if (p == 0 || (kk[p] >> rbits) != (kk[p - 1] >> rbits))
          s_cnt[kk[p] >> rbits] = rank;
        ++rank;
      }
    }
    __syncthreads();
    rank -= cnt;
    for (int u = 0; u < kPer; ++u) {
      if (!((flags >> u) & 1u)) continue;
      const uint32_t p = tid * kPer + u;
      const uint64_t key = kk[p];
      const uint32_t q = (uint32_t)(key >> rbits);
      double sum = s_val[pp[p]];
      for (uint32_t r = p + 1; r < np && kk[r] == key; ++r) sum += s_val[pp[r]];
      const uint64_t s = s_seg[q];
      const uint64_t dst = A.out_off[s] + (rank - s_cnt[q]);
      if (dst < A.capacity) {
        A.ck[dst] = (key & ((uint64_t(1) << rbits) - 1)) + A.colkey[A.seg_first[s]];
        A.cv[dst] = sum;
      }
      ++rank;
    }
    __syncthreads();
  }
}

// First index x in [c, end) with ak[x] >= target (ak ascending), probing
// c, c+1, c+3, c+7, ... before a binary search: the answer is usually a few
// entries past the cursor.
__device__ inline uint64_t gallop(const uint64_t *ak, uint64_t c, uint64_t end, uint64_t target) {
  uint64_t lo = c, hi = c, step = 1;
  while (hi < end && ak[hi] < target) {
    lo = hi + 1;
    hi = hi + step < end ? hi + step : end;
    step <<= 1;
  }
  return lower_bound_u64(ak, lo, hi, target);
}

// ----------------------------------------------- 4/5. large segments, windows
// One segment per work item.  Rows are taken in windows [r0, r1) holding at
// most kCap products (about `window` aimed for); every entry's share of a
// window is found by galloping from its cursor; empty stretches are skipped
// by jumping to the smallest next row.  Symbolic: distinct rows counted in an
// LDS hash set.  Numeric: each window is sorted by (row, position) and its
// runs summed, so rows are summed in ascending B-row order.
template <bool NUMERIC, typename K>
__global__ __launch_bounds__(kBlock) void batch_large_kernel(ColumnArgs A, const uint32_t *__restrict__ list3,
                                                             uint64_t *__restrict__ cur, uint32_t *__restrict__ wcnt,
                                                             uint32_t *__restrict__ woff, uint32_t window,
                                                             uint64_t *__restrict__ ctr, uint32_t count_slot) {
  // K: key word of the window sort (window-relative rows); 32 bits when
  // every instance's rows fit, which halves the sort's LDS.
  using HK = typename std::conditional<sizeof(K) == 4, unsigned int, unsigned long long>::type;
  __shared__ HK s_hash[NUMERIC ? 1 : kHashSlots];
  __shared__ K s_key[NUMERIC ? kCap : 1];
  __shared__ uint16_t s_pos[NUMERIC ? kCap : 1];
  __shared__ K s_key2[NUMERIC ? kCap : 1];
  __shared__ uint16_t s_pos2[NUMERIC ? kCap : 1];
  __shared__ uint32_t s_rcnt[NUMERIC ? 4 * 256 : 1];
  __shared__ double s_val[NUMERIC ? kCap : 1];
  __shared__ uint64_t s_red[kBlock / kWave];
  __shared__ double s_fred[kBlock / kWave];
  __shared__ uint32_t s_scan[kBlock / kWave];
  __shared__ uint64_t s_item;
  const uint32_t tid = threadIdx.x;
  const uint64_t nseg = ctr[kCtrNseg];
  const uint64_t n3 = ctr[count_slot];  // segments in list3
  for (;;) {
    if (tid == 0) s_item = atomicAdd((unsigned long long *)&ctr[kCtrTicketLarge], 1ull);
    __syncthreads();
    const uint64_t item = s_item;
    __syncthreads();
    if (item >= n3) break;
    const uint64_t s = list3[item];
    const uint64_t first = A.seg_first[s], last = seg_last(A.seg_first, s, nseg, A.nb);
    const uint64_t np = A.prod_off[last] - A.prod_off[first];
    const uint64_t h = A.inst[A.einst[first]].h;
    const uint64_t ckey = A.colkey[first];
    const uint64_t obase = NUMERIC ? A.out_off[s] : 0;
    for (uint64_t e = first + tid; e < last; e += kBlock) cur[e] = A.a_lo[e];
    __syncthreads();
    uint64_t done = 0, written = 0, r0 = 0;
    while (done < np) {
      const uint64_t left = np - done, span = h - r0;
      uint64_t S = span;
      if (left > window) {
        double est = (double)span * (double)window / (double)left;
        S = est < 1.0 ? 1 : (est >= (double)span ? span : (uint64_t)est);
      }
      uint64_t cnt, r1;
      for (;;) {
        r1 = S >= span ? h : r0 + S;
        uint64_t local = 0;
        for (uint64_t e = first + tid; e < last; e += kBlock) {
          const uint64_t c = cur[e], end = A.a_lo[e] + A.a_len[e];
          const uint64_t x = gallop(A.ak, c, end, A.a_base[e] + r1);
          wcnt[e] = (uint32_t)(x - c);
          local += x - c;
        }
        cnt = block_sum_u64(local, s_red);
        if (cnt <= kCap || S == 1) break;
        uint64_t S2 = (uint64_t)((double)S * (double)window / (double)cnt);
        S = S2 < 1 ? 1 : (S2 < S ? S2 : S / 2);
      }
      if (cnt == 0) {
        uint64_t local = ~0ull;
        for (uint64_t e = first + tid; e < last; e += kBlock) {
          const uint64_t c = cur[e], end = A.a_lo[e] + A.a_len[e];
          if (c < end) {
            const uint64_t row = A.ak[c] - A.a_base[e];
            local = row < local ? row : local;
          }
        }
        r0 = block_min_u64(local, s_red);
        if (r0 >= h) break;  // unreachable while the cursors agree with np
        continue;
      }
      if (cnt > kCap) {
        // One row (r1 = r0 + 1) met by more than kCap products, each from a
        // different A column: a tree sum over the entries.
        double part = 0.0;
        if (NUMERIC)
          for (uint64_t e = first + tid; e < last; e += kBlock)
            if (wcnt[e]) part += A.bv[e] * A.av[cur[e]];
        const double sum = block_sum_f64(part, s_fred);
        if (NUMERIC && tid == 0 && obase + written < A.capacity) {
          A.ck[obase + written] = r0 + ckey;
          A.cv[obase + written] = sum;
        }
        for (uint64_t e = first + tid; e < last; e += kBlock) cur[e] += wcnt[e];
        __syncthreads();
        done += cnt;
        written += 1;
        r0 = r1;
        continue;
      }
      // Window offsets of every entry (chunked block scan over the entries).
      uint32_t running = 0;
      for (uint64_t c0 = first; c0 < last; c0 += kBlock) {
        const uint64_t e = c0 + tid;
        const uint32_t v = e < last ? wcnt[e] : 0;
        uint32_t tot;
        const uint32_t ex = block256_exclusive_sum(v, s_scan, &tot);
        if (e < last) woff[e] = running + ex;
        running += tot;
      }
      if (!NUMERIC)
        for (uint32_t i = tid; i < kHashSlots; i += kBlock) s_hash[i] = (HK)kEmptyKey;
      __syncthreads();
      const uint32_t n = (uint32_t)cnt;
      uint32_t fresh = 0;
      uint64_t maxrel = 0;
      const uint32_t p0 = tid * kPer;
      if (p0 < n) {
        uint64_t e = upper_index(woff, first, last, p0);
        uint64_t wend = e + 1 < last ? woff[e + 1] : n;
        uint64_t c = cur[e] - woff[e], ab = A.a_base[e];
        double bval = NUMERIC ? A.bv[e] : 0.0;
        for (int u = 0; u < kPer; ++u) {
          const uint32_t p = p0 + u;
          if (p >= n) break;
          if (p >= wend) {
            e = upper_index(woff, e + 1, last, p);
            wend = e + 1 < last ? woff[e + 1] : n;
            c = cur[e] - woff[e];
            ab = A.a_base[e];
            if (NUMERIC) bval = A.bv[e];
          }
          const uint64_t x = c + p;
          const uint64_t row = A.ak[x] - ab;
          if (NUMERIC) {
            const uint64_t rel = row - r0;  // rows of the window lie in [r0, r1)
            maxrel = rel > maxrel ? rel : maxrel;
            s_key[p] = (K)rel;
            s_pos[p] = (uint16_t)p;
            s_val[p] = bval * A.av[x];
          } else {
            fresh += hash_insert(s_hash, row);
          }
        }
      }
      uint32_t heads;
      if (!NUMERIC) {
        block256_exclusive_sum(fresh, s_scan, &heads);
      } else {
        const uint32_t relbits = bit_width(block_max_u64(maxrel, s_red));
        const int which = lds_radix_sort<K>(s_key, s_pos, s_key2, s_pos2, n, relbits, 0, 0, s_rcnt, s_scan);
        const K *kk = which ? s_key2 : s_key;
        const uint16_t *pp = which ? s_pos2 : s_pos;
        uint32_t flags = 0, hc = 0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
          const uint32_t p = tid * kPer + u;
          const bool hd = p < n && (p == 0 || kk[p] != kk[p - 1]);
          flags |= (hd ? 1u : 0u) << u;
          hc += hd;
        }
        uint32_t rank = block256_exclusive_sum(hc, s_scan, &heads);
        for (int u = 0; u < kPer; ++u) {
          if (!((flags >> u) & 1u)) continue;
          const uint32_t p = tid * kPer + u;
          const uint64_t key = kk[p];
          double sum = s_val[pp[p]];
          for (uint32_t r = p + 1; r < n && kk[r] == key; ++r) sum += s_val[pp[r]];
          const uint64_t dst = obase + written + rank;
          if (dst < A.capacity) {
            A.ck[dst] = r0 + key + ckey;
            A.cv[dst] = sum;
          }
          ++rank;
        }
      }
      for (uint64_t e = first + tid; e < last; e += kBlock) cur[e] += wcnt[e];
      __syncthreads();
      done += cnt;
      written += heads;
      r0 = r1;
    }
    if (!NUMERIC && tid == 0) A.seg_u[s] = written;
    __syncthreads();
  }
}

// ------------------------------------------ 4/5. large segments, row pieces
// A large segment of at most kPieceLists B entries is cut by row value into
// q equal row ranges ("pieces") of about kPieceTarget products: every A
// column is sorted by row, so a piece is one contiguous run of each column.
// Pieces are independent work items: one 512-thread
// workgroup finds its runs (four lanes per bound, in rounds of four probes),
// loads their rows into LDS, merges them pairwise (ceil(log2 c) merge-path
// rounds, the earlier B entry first on equal rows) and sums equal rows in
// that order -- the chain of additions, in ascending B-row order, of the
// window sort.  The symbolic pass counts the piece's distinct rows (into the
// piece's slot and its segment's count); the numeric pass takes its output
// offset from its predecessors' counts and writes its rows.  A piece of more
// than kPieceCap products (rows clustered in one range) goes through in row
// sub-ranges found by halving.  Piece slots come from a pool of fixed size;
// a segment that does not fit in it, has more than kPieceLists entries, or
// spans more than 2^32 rows per piece takes the windows instead.
constexpr uint32_t kPieceBlock = 512;
constexpr uint32_t kPieceCap = 3584;  // 7 per thread; 49 KB of LDS: three workgroups per CU
constexpr uint32_t kPieceTarget = 3072;
constexpr uint32_t kPieceLists = 64;
constexpr uint32_t kPieceItems = kPieceCap / kPieceBlock;
constexpr uint32_t kNoPiece = 0xFFFFFFFFu;

struct PieceArgs {
  uint64_t *rmin, *width, *pbase;  // per large segment (position in list3)
  uint32_t *pq;                    // its pieces; 0: the windows take it
  uint32_t *piece_seg, *piece_u;   // per piece slot: position in list3 (kNoPiece: unused), distinct rows
  uint64_t *piece_ent;             // per piece slot: first B entry << 8 | entries
  uint64_t *status;                // per piece slot: look-back word of the direct pass
  uint32_t *list3w;                // the segments the windows take
  uint32_t *bnd;                   // per piece slot: kPieceBnd first A entries of its runs (offsets in
                                   // their columns), for pieces of at most kPieceBnd B entries
  uint64_t cap;                    // piece slots
};

// Pieces of at most kPieceBnd B entries find their runs in bnd (searched
// beforehand, all pieces at once) instead of searching the columns first.
constexpr uint32_t kPieceBnd = 16;
constexpr size_t kPieceSlotBytes = 24 + 4 * kPieceBnd;

// Per large segment (one wave each, lane i on B entry i): the row range of
// its A columns, the piece width and the number of pieces (0: the windows
// take it).  pq is written for all nb slots (zero past the last large
// segment), so that its scan gives every segment its first piece slot and
// the total the batch needs.
__global__ __launch_bounds__(kBlock) void batch_piece_plan_kernel(ColumnArgs A, const uint32_t *__restrict__ list3,
                                                                  PieceArgs Q, const uint64_t *__restrict__ ctr) {
  const uint64_t n3 = ctr[kCtrN3], nseg = ctr[kCtrNseg];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  for (uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < A.nb; r += waves) {
    if (r >= n3) {
      if (lane == 0) Q.pq[r] = 0;
      continue;
    }
    const uint64_t s = list3[r];
    const uint64_t first = A.seg_first[s], last = seg_last(A.seg_first, s, nseg, A.nb);
    const uint64_t c = last - first, np = A.prod_off[last] - A.prod_off[first];
    uint64_t q = 0, lo = ~0ull, hi = 0, w = 0;
    if (c <= kPieceLists) {
      if (lane < c) {
        const uint64_t e = first + lane;
        const uint32_t n = A.a_len[e];
        if (n) {
          lo = A.ak[A.a_lo[e]] - A.a_base[e];
          hi = A.ak[A.a_lo[e] + n - 1] - A.a_base[e];
        }
      }
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint64_t ol = __shfl_xor(lo, off, kWave), oh = __shfl_xor(hi, off, kWave);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
      }
      // np > 0 (a large segment), so some column is non-empty and lo <= hi.
      const uint64_t range = hi - lo + 1;
      q = (np + kPieceTarget - 1) / kPieceTarget;
      w = (range + q - 1) / q;
      if (w > 0xFFFFFFFFull) q = 0;
      else q = (range + w - 1) / w;
    }
    if (lane == 0) {
      Q.rmin[r] = lo;
      Q.width[r] = w;
      Q.pq[r] = (uint32_t)q;
    }
  }
}

// After the scan of pq into pbase: the piece slots of every segment that
// fits in the pool (the first ones, as slots ascend with the segment's
// position in list3); the others go to the windows.
__global__ __launch_bounds__(kBlock) void batch_piece_assign_kernel(ColumnArgs A, const uint32_t *__restrict__ list3,
                                                                    PieceArgs Q, uint64_t *__restrict__ ctr) {
  const uint64_t n3 = ctr[kCtrN3], nseg = ctr[kCtrNseg];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  for (uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < n3; r += waves) {
    const uint64_t q = Q.pq[r], base = Q.pbase[r];
    const bool fits = q && base + q <= Q.cap;
    if (fits) {
      const uint64_t s = list3[r], first = A.seg_first[s], last = seg_last(A.seg_first, s, nseg, A.nb);
      for (uint64_t t = base + lane; t < base + q; t += kWave) {
        Q.piece_seg[t] = (uint32_t)r;
        Q.piece_ent[t] = first << 8 | (last - first);
      }
    }
    else if (q && base < Q.cap)  // the one segment that straddles the pool's end
      for (uint64_t t = base + lane; t < Q.cap; t += kWave) Q.piece_seg[t] = kNoPiece;
    if (lane == 0 && !fits) {
      Q.pq[r] = 0;
      Q.list3w[atomicAdd((unsigned long long *)&ctr[kCtrN3w], 1ull)] = list3[r];
    }
  }
}

// bnd[t * kPieceBnd + i]: the first A entry of piece slot t in the column of
// its segment's B entry i (an offset in that column), for every piece of at
// most kPieceBnd entries -- one thread per (piece, entry): the searches run
// side by side instead of on each piece's path.
__global__ __launch_bounds__(kBlock) void batch_piece_bounds_kernel(ColumnArgs A, PieceArgs Q,
                                                                    const uint64_t *__restrict__ ctr) {
  const uint64_t npieces = min<uint64_t>(ctr[kCtrPieces], Q.cap);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < npieces * kPieceBnd; x += stride) {
    const uint64_t t = x / kPieceBnd;
    const uint32_t i = (uint32_t)(x % kPieceBnd);
    const uint32_t r = Q.piece_seg[t];
    if (r == kNoPiece) continue;
    const uint64_t ent = Q.piece_ent[t];
    const uint32_t c = (uint32_t)(ent & 0xFF);
    if (c > kPieceBnd || i >= c) continue;
    const uint64_t e = (ent >> 8) + i, p = t - Q.pbase[r];
    const uint64_t base = A.a_lo[e];
    uint64_t lo = base, hi = base + A.a_len[e];
    if (p > 0) {
      const uint64_t key = A.a_base[e] + Q.rmin[r] + p * Q.width[r];
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (A.ak[mid] < key) lo = mid + 1;
        else hi = mid;
      }
    }
    Q.bnd[x] = (uint32_t)(lo - base);
  }
}

// Status word of a piece in the direct pass: [63:62] 1 = its own row count,
// 2 = the count of every piece up to and including it; [61:0] the count.
constexpr uint64_t kPieceAgg = uint64_t(1) << 62;
constexpr uint64_t kPieceIncl = uint64_t(2) << 62;
constexpr uint64_t kPieceValue = kPieceAgg - 1;

// NUMERIC false: the symbolic pass (row counts).  NUMERIC true: the numeric
// pass, at offsets from the symbolic pass's counts; with DIRECT (a recorded
// graph, whose output arrays are known before any count), the symbolic pass
// skips the pieces and this pass chains the pieces' counts by a decoupled
// look-back over all pieces in output order (pieces take tickets, so a piece
// waits only on pieces already running), adding them to the offsets of the
// other segments; piece_total then records each pieced segment's count.
template <bool NUMERIC, bool DIRECT>
__global__ __launch_bounds__(kPieceBlock, 6) void batch_piece_kernel(ColumnArgs A, const uint32_t *__restrict__ list3,
                                                                     PieceArgs Q, uint64_t *__restrict__ ctr) {
  static_assert(NUMERIC || !DIRECT, "the direct pass writes C");
  // Rows (offsets from the piece's first row) and load positions, in two
  // buffers for the merge rounds, and the B entry of each load position.
  __shared__ uint32_t s_row[2][kPieceCap];
  __shared__ uint16_t s_idx[2][kPieceCap];
  __shared__ uint8_t s_run[kPieceCap];
  __shared__ uint32_t s_off[kPieceLists + 1];
  __shared__ uint64_t s_src[kPieceLists];   // first A entry of each run of the current range
  __shared__ uint64_t s_next[kPieceLists];  // first A entry past the current range
  __shared__ uint64_t s_end[kPieceLists];   // first A entry past the piece
  __shared__ uint64_t s_key0[kPieceLists];  // key of the piece's first row in each column
  __shared__ double s_b[kPieceLists];
  __shared__ uint64_t s_scan[kPieceBlock / kWave];
  __shared__ uint64_t s_all, s_lo, s_base, s_t, s_prefix;
  const uint32_t tid = threadIdx.x;
  const uint64_t nseg = ctr[kCtrNseg];
  // Without DIRECT, pieces go in a fixed stride over the persistent grid (no
  // shared ticket: pieces are alike in size, and one contended counter costs
  // a little on every piece).
  const uint64_t npieces = min<uint64_t>(ctr[kCtrPieces], Q.cap);
  uint64_t t_next = blockIdx.x;
  for (;;) {
    uint64_t t;
    if (DIRECT) {
      if (tid == 0) s_t = atomicAdd((unsigned long long *)&ctr[kCtrTicketPiece], 1ull);
      __syncthreads();
      t = s_t;
      __syncthreads();
    } else {
      t = t_next;
      t_next += gridDim.x;
    }
    if (t >= npieces) break;
    // Loads that do not wait on one another: the piece's segment position
    // and entries (both before the test of the first), then the segment's
    // plan and the entries' columns.
    const uint32_t r = Q.piece_seg[t];
    const uint64_t ent = Q.piece_ent[t];
    if (r == kNoPiece) {  // a slot past the pool's last fitting segment: no rows
      if (DIRECT && tid == 0) status_store(&Q.status[t], t == 0 ? kPieceIncl : kPieceAgg);
      continue;
    }
    const uint64_t first = ent >> 8;
    const uint32_t c = (uint32_t)(ent & 0xFF);
    const uint64_t s = list3[r];
    const uint64_t q = Q.pq[r], p = t - Q.pbase[r], w = Q.width[r], v0 = Q.rmin[r] + p * w;
    const bool last_piece = p + 1 == q;
    if (DIRECT && tid == 0) s_base = A.out_off[s];  // plus the look-back's prefix
    if (NUMERIC && !DIRECT && tid < kWave) {
      // Output offset: the segment's, plus the rows of its earlier pieces
      // (issued before the searches, which do not need it).
      uint64_t before = 0;
      const uint64_t b0 = Q.pbase[r];
      for (uint64_t j = tid; j < p; j += kWave) before += Q.piece_u[b0 + j];
      before = __shfl(wave_inclusive_sum64(before), kWave - 1, kWave);
      if (tid == 0) s_base = A.out_off[s] + before;
    }
    if (c <= kPieceBnd) {
      // The piece's runs from the bounds searched beforehand: its own start
      // and the next piece's (the column's end for the last piece).
      if (tid < c) {
        const uint64_t e = first + tid, lo = A.a_lo[e];
        s_src[tid] = lo + Q.bnd[t * kPieceBnd + tid];
        s_end[tid] = last_piece ? lo + A.a_len[e] : lo + Q.bnd[(t + 1) * kPieceBnd + tid];
        s_key0[tid] = A.a_base[e] + v0;
        s_b[tid] = A.bv[e];
      }
    } else {
      // The piece's runs: 2G lanes per B entry, G per side (0: its first A
      // entry in the piece, 1: the first past it -- the column's end for the
      // last piece), G the largest power of two up to 32 that fits all
      // entries in the block.
      uint32_t G = 32;
      while (G > 4 && 2 * G * c > kPieceBlock) G >>= 1;
      const uint32_t i = tid / (2 * G), side = (tid / G) & 1;
      uint64_t lo = 0, hi = 0, key = 0;
      if (i < c) {
        const uint64_t e = first + i;
        lo = A.a_lo[e];
        hi = lo + A.a_len[e];
        key = A.a_base[e] + v0 + (side ? w : 0);
      }
      const bool search = i < c && !(side ? last_piece : p == 0);
      const uint64_t bnd = group_lower_bound(A.ak, lo, hi, key, search, G);
      if (i < c && (tid & (G - 1)) == 0) {
        const uint64_t e = first + i;
        if (side == 0) {
          s_src[i] = bnd;
          s_key0[i] = A.a_base[e] + v0;
          s_b[i] = A.bv[e];
        } else {
          s_end[i] = last_piece ? hi : bnd;
        }
      }
    }
    __syncthreads();
    if (tid < kWave) {  // c <= kPieceLists = one wave: lane i takes run i
      uint64_t len = 0;
      if (tid < c) {
        len = s_end[tid] - s_src[tid];
        s_next[tid] = s_end[tid];
      }
      const uint64_t incl = wave_inclusive_sum64(len);
      if (tid < c) s_off[tid] = (uint32_t)min<uint64_t>(incl - len, kPieceCap + 1);
      if (tid == kWave - 1) {
        s_off[c] = (uint32_t)min<uint64_t>(incl, kPieceCap + 1);
        s_all = incl;
      }
    }
    __syncthreads();
    const uint64_t total_all = s_all;
    const uint64_t ckey = A.colkey[first];
    auto run_of = [&](uint32_t x) __attribute__((always_inline)) {
      uint32_t lo = 0, hi = c;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= x) lo = mid;
        else hi = mid;
      }
      return lo;
    };
    uint32_t cur = 0;
    // Loads the runs s_src[i] + [0, s_off[i+1] - s_off[i]) (at most
    // kPieceCap products), orders them by (row, B entry), lists the first
    // position of every distinct row and returns their number; `cur` names
    // the ordered buffer, s_row[cur ^ 1] holds the list.
    auto order_rows = [&]() __attribute__((always_inline)) -> uint64_t {
      const uint32_t total = s_off[c];
      for (uint32_t x = tid; x < total; x += kPieceBlock) {
        const uint32_t i = run_of(x);
        const uint64_t ai = s_src[i] + (x - s_off[i]);
        s_row[0][x] = (uint32_t)(A.ak[ai] - s_key0[i]);
        s_idx[0][x] = (uint16_t)x;
        s_run[x] = (uint8_t)i;
      }
      __syncthreads();
      // Pairwise merge rounds: after the round of width mw, every run holds
      // 2mw consecutive B entries.  Each thread writes one contiguous
      // stretch of output positions, located in its pair by a merge-path
      // search; on equal rows the left run (earlier B entries) goes first.
      cur = 0;
      const uint32_t per = (total + kPieceBlock - 1) / kPieceBlock;
      for (uint32_t mw = 1; mw < c; mw <<= 1) {
        const uint32_t *kin = s_row[cur];
        const uint16_t *iin = s_idx[cur];
        uint32_t *kout = s_row[cur ^ 1];
        uint16_t *iout = s_idx[cur ^ 1];
        uint32_t o = tid * per;
        const uint32_t oe = min(o + per, total);
        while (o < oe) {
          const uint32_t g = run_of(o) / (2 * mw) * (2 * mw);
          const uint32_t a0 = s_off[g], a1 = s_off[min(c, g + mw)], b1 = s_off[min(c, g + 2 * mw)];
          const uint32_t na = a1 - a0, nbr = b1 - a1, d = o - a0;
          uint32_t lo = d > nbr ? d - nbr : 0, hi = d < na ? d : na;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (kin[a0 + mid] <= kin[a1 + d - 1 - mid]) lo = mid + 1;
            else hi = mid;
          }
          uint32_t i = a0 + lo, j = a1 + (d - lo);
          const uint32_t stop = min(oe, b1);
          uint32_t ka = i < a1 ? kin[i] : ~0u, kb = j < b1 ? kin[j] : ~0u;
          for (; o < stop; ++o) {
            if (j >= b1 || (i < a1 && ka <= kb)) {
              kout[o] = ka;
              iout[o] = iin[i];
              ++i;
              ka = i < a1 ? kin[i] : ~0u;
            } else {
              kout[o] = kb;
              iout[o] = iin[j];
              ++j;
              kb = j < b1 ? kin[j] : ~0u;
            }
          }
        }
        __syncthreads();
        cur ^= 1;
      }
      // Distinct rows over `per` consecutive positions per thread.
      const uint32_t *krow = s_row[cur];
      uint32_t flags = 0, cnt = 0;
      const uint32_t h0 = tid * per, h1 = min(h0 + per, total);
      for (uint32_t x = h0; x < h1; ++x) {
        const bool h = x == 0 || krow[x] != krow[x - 1];
        flags |= (h ? 1u : 0u) << (x - h0);
        cnt += h;
      }
      uint64_t utotal;
      uint64_t urank = block_exclusive_sum_waves<kPieceBlock / kWave>(cnt, s_scan, &utotal);
      if (NUMERIC) {
        uint32_t *s_head = s_row[cur ^ 1];  // free after the last merge round
        for (uint32_t x = h0; x < h1; ++x)
          if ((flags >> (x - h0)) & 1u) s_head[urank++] = x;
      }
      __syncthreads();
      return utotal;
    };
    // The ordered rows' sums, kept in registers (row u = tid + k *
    // kPieceBlock in slot k), so that a direct piece gathers them before its
    // look-back waits; the rows themselves stay in LDS (the head list and
    // the ordered rows), read again when they are stored.
    double hsum[kPieceItems];
    auto sum_rows = [&](uint64_t utotal) __attribute__((always_inline)) {
      // Each product rounded on its own, never fused with the addition that
      // follows (as in the window sort).  The first product of every row is
      // fetched for all of a thread's rows before any is summed.
#pragma clang fp contract(off)
      const uint32_t total = s_off[c];
      const uint32_t *krow = s_row[cur], *s_head = s_row[cur ^ 1];
      const uint16_t *kidx = s_idx[cur];
      auto product = [&](uint32_t y) __attribute__((always_inline)) {
        const uint32_t x = kidx[y], i = s_run[x];
        return s_b[i] * A.av[s_src[i] + (x - s_off[i])];
      };
      const uint32_t nu = (uint32_t)((utotal + kPieceBlock - 1) / kPieceBlock);
#pragma unroll
      for (uint32_t k = 0; k < kPieceItems; ++k) {
        const uint32_t u = tid + k * kPieceBlock;
        if (k < nu && u < utotal) hsum[k] = product(s_head[u]);
      }
#pragma unroll
      for (uint32_t k = 0; k < kPieceItems; ++k) {
        const uint32_t u = tid + k * kPieceBlock;
        if (k < nu && u < utotal) {
          const uint32_t x = s_head[u], row = krow[x];
          for (uint32_t y = x + 1; y < total && krow[y] == row; ++y) hsum[k] += product(y);
        }
      }
    };
    // Stores the summed rows at dst0 of C.
    auto store_rows = [&](uint64_t dst0, uint64_t utotal) __attribute__((always_inline)) {
      const uint32_t nu = (uint32_t)((utotal + kPieceBlock - 1) / kPieceBlock);
      const uint32_t *krow = s_row[cur], *s_head = s_row[cur ^ 1];
#pragma unroll
      for (uint32_t k = 0; k < kPieceItems; ++k) {
        const uint32_t u = tid + k * kPieceBlock;
        if (k < nu && u < utotal && dst0 + u < A.capacity) {
          A.ck[dst0 + u] = v0 + krow[s_head[u]] + ckey;
          A.cv[dst0 + u] = hsum[k];
        }
      }
    };
    // Direct pass: wave 0 sums its predecessors' words, 64 per round trip,
    // back to the first inclusive one, into s_prefix.  Only counts travel
    // between pieces (C is read by later launches), so relaxed atomics
    // suffice.
    auto look_back = [&]() __attribute__((always_inline)) {
      if (tid < kWave) {
        uint64_t prefix = 0;
        if (t > 0) {
          int64_t base = (int64_t)t - 1;
          for (;;) {
            const int64_t idx = base - (int64_t)tid;
            uint64_t wd = 0;
            bool ready = true, term = true;
            if (idx >= 0) {
              wd = status_load(&Q.status[idx]);
              ready = wd != 0;
              term = ready && (wd & kPieceIncl);
            }
            const uint64_t term_mask = __ballot(term), wait_mask = __ballot(!ready);
            const uint32_t stop = term_mask ? (uint32_t)__builtin_ctzll(term_mask) : kWave - 1;
            const uint64_t needed = stop == kWave - 1 ? ~0ull : ((2ull << stop) - 1ull);
            if (wait_mask & needed) {
              __builtin_amdgcn_s_sleep(1);
              continue;
            }
            const uint64_t part = tid <= stop ? (wd & kPieceValue) : 0;
            prefix += __shfl(wave_inclusive_sum64(part), kWave - 1, kWave);
            if (term_mask) break;
            base -= kWave;
          }
        }
        if (tid == 0) s_prefix = prefix;
      }
      __syncthreads();
    };
    uint64_t count = 0;
    if (total_all <= kPieceCap) {
      count = order_rows();
      if (DIRECT) {
        if (tid == 0) status_store(&Q.status[t], (t == 0 ? kPieceIncl : kPieceAgg) | count);
        sum_rows(count);
        look_back();
        if (tid == 0 && t > 0) status_store(&Q.status[t], kPieceIncl | (s_prefix + count));
        store_rows(s_base + s_prefix, count);
      } else if (NUMERIC) {
        sum_rows(count);
        store_rows(s_base, count);
      }
    } else {
      // An oversized piece cannot publish its count early: it takes its
      // prefix first, and its successors wait for its inclusive count.
      if (DIRECT) look_back();
      const uint64_t dst = DIRECT ? s_base + s_prefix : s_base;
      // Rows clustered: row sub-ranges of at most kPieceCap products, each
      // found by halving the remaining range (one row holds at most c <=
      // kPieceLists products, so the halving ends), taken one after another.
      if (tid == 0) s_lo = 0;
      __syncthreads();
      for (;;) {
        // Every thread reads the rows done before wave 0 moves them on.
        const uint64_t done_rows = s_lo;
        if (done_rows >= w) break;
        __syncthreads();
        if (tid < kWave) {
          uint64_t span = w - done_rows, nxt = 0;
          for (;;) {
            uint64_t len = 0;
            if (tid < c) {  // first A entry of run tid at or past row s_lo + span
              const uint64_t key = s_key0[tid] + done_rows + span;
              uint64_t lo = s_src[tid], hi = s_end[tid];
              while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (A.ak[mid] < key) lo = mid + 1;
                else hi = mid;
              }
              nxt = lo;
              len = lo - s_src[tid];
            }
            const uint64_t tot = __shfl(wave_inclusive_sum64(len), kWave - 1, kWave);
            if (tot <= kPieceCap || span == 1) break;
            span = (span + 1) / 2;
          }
          const uint64_t len = tid < c ? nxt - s_src[tid] : 0;
          const uint64_t incl = wave_inclusive_sum64(len);
          if (tid < c) {
            s_off[tid] = (uint32_t)(incl - len);
            s_next[tid] = nxt;
          }
          if (tid == kWave - 1) s_off[c] = (uint32_t)incl;
          if (tid == 0) s_lo = done_rows + span;
        }
        __syncthreads();
        const uint64_t u = order_rows();
        if (NUMERIC) {
          sum_rows(u);
          store_rows(dst + count, u);
        }
        count += u;
        __syncthreads();
        if (tid < c) s_src[tid] = s_next[tid];
        __syncthreads();
      }
    }
    if (DIRECT && total_all > kPieceCap && tid == 0) status_store(&Q.status[t], kPieceIncl | (s_prefix + count));
    if (!NUMERIC && tid == 0) {
      Q.piece_u[t] = (uint32_t)count;
      atomicAdd((unsigned long long *)&A.seg_u[s], (unsigned long long)count);
    }
    __syncthreads();  // LDS is reused by the next piece
  }
}

// Direct pass: each pieced segment's row count, the difference of the
// inclusive counts that close it and its predecessor (every slot before a
// fitting segment's first piece holds a piece of an earlier one).
__global__ void batch_piece_total_kernel(const uint32_t *__restrict__ list3, PieceArgs Q, uint64_t *__restrict__ seg_u,
                                         const uint64_t *__restrict__ ctr) {
  const uint64_t n3 = ctr[kCtrN3];
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n3; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q = Q.pq[r];
    if (!q) continue;
    const uint64_t b0 = Q.pbase[r];
    const uint64_t hi = Q.status[b0 + q - 1] & kPieceValue, lo = b0 ? Q.status[b0 - 1] & kPieceValue : 0;
    seg_u[list3[r]] = hi - lo;
  }
}

// --------------------------------- 4/5. dense segments (chosen per instance)
// An instance whose accumulator is dense (lnt_batch_set_accumulators) and
// whose lattice height h is at most kDenseH sends its segments of two or more
// entries and at least dense_min products (8 or more per entry on average)
// here: one workgroup per segment, a presence flag byte per row and, in the
// numeric pass, h doubles of accumulator in LDS, indexed by the row -- the
// sparse accumulator of csc_times (SparseLatticeBase.h:2224-2351).  Flags are
// bytes, set by plain stores: a bitmap set by LDS atomics serialises the
// lanes of a wave that hit one word, which dense columns do all the time.  The
// segment is taken in chunks of at most BT*U products over at most kDenseEnt
// entries: every load of a chunk is issued at once, then (numeric) its entries
// are applied one after another with a barrier between them (the rows of one
// entry are distinct, so its threads never meet on a row).  A row's first
// product is stored and the later ones added in entry order, which is
// ascending B-row order: the same chain of additions as the sort paths, so
// the values are bit-identical to them.  Symbolic: the number of flags set is
// the segment's output count.  Numeric: the flags are packed into a bitmap,
// and output position o (consecutive threads, consecutive positions) finds
// its bitmap word by a search over the words' ranks and its row by selecting
// the bit.
//
// The work is latency-bound (a descriptor, the chunk's entry tables and its
// products are three dependent loads), so the kernels keep many segments in
// flight: the symbolic pass (flags only) runs 256 threads per segment; the
// numeric pass has a 16 KB accumulator variant for heights up to
// kDenseSmallH (eight workgroups per compute unit) and a 64 KB one for the
// rest (two).  Workgroups take segments from a shared counter a few at a
// time (TicketBlock).
//   which = 0: symbolic, every dense segment; 1: numeric, heights up to
//   kDenseSmallH; 2: numeric, taller (descriptors from the back of `desc`).

// Position of the k-th (0-based) set bit of x; x has more than k set bits.
__host__ __device__ inline uint32_t select_bit(uint32_t x, uint32_t k) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t width = 16; width >= 1; width >>= 1) {
    const uint32_t lo = x & ((1u << width) - 1u);
    const uint32_t c = (uint32_t)__builtin_popcount(lo);
    if (k >= c) {
      k -= c;
      x >>= width;
      pos += width;
    } else {
      x = lo;
    }
  }
  return pos;
}

// Four flag bytes (each 0 or 1) of a word -> four bits.
__device__ inline uint32_t flag_nibble(uint32_t f) { return (f | (f >> 7) | (f >> 14) | (f >> 21)) & 0xFu; }

template <bool NUMERIC, int BT, int U, uint32_t HMAX>
__global__ __launch_bounds__(BT) void batch_dense_kernel(ColumnArgs A, const DenseDesc *__restrict__ desc, int which,
                                                         uint64_t *__restrict__ ctr) {
  constexpr uint32_t kWords = HMAX / 32;
  constexpr uint32_t kWpt = (kWords + BT - 1) / BT;  // bitmap words per thread
  constexpr uint32_t kChunk = (uint32_t)BT * U;
  __shared__ double s_acc[NUMERIC ? HMAX : 1];
  __shared__ uint32_t s_flag32[HMAX / 4];  // one flag byte per row
  __shared__ uint32_t s_bits[NUMERIC ? kWords : 1];
  __shared__ uint32_t s_wrank[NUMERIC ? kWords : 1];
  __shared__ uint64_t s_poff[kDenseEnt + 1];
  __shared__ uint64_t s_alo[kDenseEnt], s_abase[kDenseEnt];
  __shared__ double s_b[NUMERIC ? kDenseEnt : 1];
  __shared__ uint32_t s_scan[BT / kWave];
  __shared__ uint64_t s_item, s_out0, s_ck0;
  const uint32_t tid = threadIdx.x;
  const uint64_t packed = ctr[kCtrNdense];
  const uint64_t nsm = packed & 0xFFFFFFFFull, nlg = packed >> 32;
  const uint64_t nitems = which == 0 ? nsm + nlg : which == 1 ? nsm : nlg;
  TicketBlock<NUMERIC ? 2 : 4> tickets;
  unsigned long long *const counter =
      (unsigned long long *)&ctr[which == 2 ? kCtrTicketDense2 : kCtrTicketDense];
  for (;;) {
    if (tid == 0) s_item = tickets.next(counter);
    __syncthreads();
    const uint64_t item = s_item;
    if (item >= nitems) break;
    const uint64_t di = which == 1 || (which == 0 && item < nsm) ? item : A.nb - (which == 0 ? item - nsm : item);
    const DenseDesc d = desc[di];
    const uint64_t first = d.first, last = d.first + d.nent;
    const uint32_t words = (d.h + 31) / 32;
    uint8_t *const s_flag = reinterpret_cast<uint8_t *>(s_flag32);
    for (uint32_t w = tid; w < 8 * words; w += BT) s_flag32[w] = 0;
    uint64_t p0 = d.p0;
    uint64_t e0 = first;  // the entry holding product p0, or an empty one before it
    while (p0 < d.p1) {
      // The previous chunk's readers of the entry tables are done (and the
      // bitmap is clear before the first chunk).
      __syncthreads();
      const uint32_t ne = (uint32_t)min<uint64_t>(kDenseEnt, last - e0);
      for (uint32_t i = tid; i <= ne; i += BT) {
        const uint64_t e = e0 + i;  // <= last: prod_off has nb + 1 entries
        s_poff[i] = A.prod_off[e];
        if (i < ne) {
          s_alo[i] = A.a_lo[e];
          s_abase[i] = A.a_base[e];
          if (NUMERIC) s_b[i] = A.bv[e];
        }
      }
      if (NUMERIC && tid == 0 && e0 == first) {
        s_out0 = A.out_off[d.seg];
        s_ck0 = A.colkey[first];
      }
      __syncthreads();
      const uint64_t p1 = min<uint64_t>(p0 + kChunk, s_poff[ne]);
      uint32_t row[U], ent[U];
      double val[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t p = p0 + tid + (uint64_t)u * BT;
        ent[u] = ~0u;
        row[u] = 0;
        val[u] = 0.0;
        if (p < p1) {
          // The entry of p: the last one whose products start at or before p
          // (empty entries share their successor's start).
          const uint32_t i = (uint32_t)upper_index(s_poff, 0, ne, p);
          const uint64_t ai = s_alo[i] + (p - s_poff[i]);
          row[u] = (uint32_t)(A.ak[ai] - s_abase[i]);
          if (NUMERIC) val[u] = s_b[i] * A.av[ai];
          ent[u] = i;
        }
      }
      if (!NUMERIC) {
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (ent[u] != ~0u) s_flag[row[u]] = 1;
      } else if (p1 > p0) {  // (a run of kDenseEnt empty entries holds no product)
        const uint32_t i_lo = (uint32_t)upper_index(s_poff, 0, ne, p0);
        const uint32_t i_hi = (uint32_t)upper_index(s_poff, 0, ne, p1 - 1);
        for (uint32_t r = i_lo; r <= i_hi; ++r) {
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (ent[u] == r) {
              if (s_flag[row[u]]) {
                s_acc[row[u]] += val[u];
              } else {
                s_acc[row[u]] = val[u];
                s_flag[row[u]] = 1;
              }
            }
          }
          __syncthreads();
        }
      }
      e0 += p1 < s_poff[ne] ? upper_index(s_poff, 0, ne, p1) : ne;
      p0 = p1;
    }
    __syncthreads();
    // Bitmap word w = the flags of rows [32w, 32w + 32) (flags past h stay 0).
    uint32_t cnt = 0, mine[kWpt];
#pragma unroll
    for (uint32_t q = 0; q < kWpt; ++q) {
      const uint32_t w = tid * kWpt + q;
      uint32_t m = 0;
      if (w < words) {
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) m |= flag_nibble(s_flag32[8 * w + j]) << (4 * j);
      }
      mine[q] = m;
      cnt += (uint32_t)__popc(m);
    }
    uint32_t total;
    uint32_t rank = block_exclusive_sum_bt<BT>(cnt, s_scan, &total);
    if (!NUMERIC) {
      if (tid == 0) A.seg_u[d.seg] = total;
    } else {
#pragma unroll
      for (uint32_t q = 0; q < kWpt; ++q) {
        const uint32_t w = tid * kWpt + q;
        if (w < words) {
          s_bits[w] = mine[q];
          s_wrank[w] = rank;
          rank += (uint32_t)__popc(mine[q]);
        }
      }
      __syncthreads();
      const uint64_t out0 = s_out0, ck0 = s_ck0;
      for (uint32_t o = tid; o < total; o += BT) {
        const uint32_t w = (uint32_t)upper_index(s_wrank, 0, words, o);
        const uint32_t r = w * 32 + select_bit(s_bits[w], o - s_wrank[w]);
        const uint64_t dst = out0 + o;
        if (dst < A.capacity) {
          A.ck[dst] = ck0 + r;
          A.cv[dst] = s_acc[r];
        }
      }
    }
    // Every thread is done with this segment's LDS before the next one.
    __syncthreads();
  }
}

// ------------------------------------------- 5. single-entry segments: copy
// C(:,j) = b * A(:,k) for every segment of one B entry.  A wave takes 64
// consecutive segments (lane j: segment g*64 + j, its A run, output offset,
// key shift and B value in registers) and spreads their products over its
// lanes: a wave prefix of the single segments' lengths numbers the products,
// a chunk of 64*kSingleItems products finds each product's segment by walking
// the few segments that meet the chunk (a ballot, then one compare per item
// and segment), takes that segment's place with lane shuffles, and issues
// every load of the chunk before its stores -- so a wave keeps that many
// loads in flight whatever the segments' lengths (they are often shorter
// than a wave).  The groups are cut into eight contiguous ranges, one per XCD
// (workgroups are dealt to the XCDs in turn): the waves of one XCD then work
// on few instances at a time, whose A columns -- each read by many segments
// -- stay in that XCD's L2 (HBM reads fell about fourfold; 18.7 -> 13.1 ms on
// the 10,000-contraction mix).  Correctness does not depend on the placement.
constexpr uint32_t kSingleItems = 8;

__device__ inline uint64_t readlane64(uint64_t v, uint32_t j) {
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, (int)j);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)j);
  return ((uint64_t)hi << 32) | lo;
}

// The grid is a multiple of 8.
__global__ __launch_bounds__(kBlock) void batch_single_kernel(ColumnArgs A, const uint32_t *__restrict__ kind,
                                                              const uint64_t *__restrict__ ctr) {
  const uint64_t nseg = ctr[kCtrNseg];
  const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint64_t ngroups = (nseg + kWave - 1) / kWave;
  const uint32_t x = blockIdx.x & 7, nx = gridDim.x >> 3;
  const uint64_t g0 = ngroups * x / 8, g1 = ngroups * (x + 1) / 8;
  const uint64_t waves = (uint64_t)nx * (kBlock / kWave);
  for (uint64_t g = g0 + (uint64_t)(blockIdx.x >> 3) * (kBlock / kWave) + w; g < g1; g += waves) {
    const uint64_t s = g * kWave + lane;
    const bool single = s < nseg && kind[s] == 1;
    if (!__ballot(single)) continue;
    uint64_t n = 0, o = 0, lo = 0, kshift = 0;
    double b = 0.0;
    if (single) {
      const uint64_t e = A.seg_first[s];
      o = A.out_off[s];
      lo = A.a_lo[e];
      n = A.a_len[e];
      kshift = A.colkey[e] - A.a_base[e];
      b = A.bv[e];
    }
    const uint64_t incl = wave_inclusive_sum64(n), excl = incl - n;
    const uint64_t ashift = lo - excl, oshift = o - excl;  // A index and output index - product number
    const uint64_t total = __shfl(incl, kWave - 1, kWave);
    for (uint64_t p0 = 0; p0 < total; p0 += (uint64_t)kWave * kSingleItems) {
      const uint64_t pend = min<uint64_t>(p0 + (uint64_t)kWave * kSingleItems, total);
      // The segments that meet [p0, pend), ascending: the first holds p0, so
      // every item gets a segment; later ones take the items past their start.
      uint64_t m = __ballot(n > 0 && excl < pend && incl > p0);
      uint32_t js[kSingleItems];
#pragma unroll
      for (uint32_t u = 0; u < kSingleItems; ++u) js[u] = 0;
      while (m) {
        const uint32_t j = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const uint64_t ej = readlane64(excl, j);
#pragma unroll
        for (uint32_t u = 0; u < kSingleItems; ++u)
          if (p0 + u * kWave + lane >= ej) js[u] = j;
      }
      uint64_t key[kSingleItems], dst[kSingleItems];
      double val[kSingleItems];
#pragma unroll
      for (uint32_t u = 0; u < kSingleItems; ++u) {
        const uint64_t p = p0 + u * kWave + lane;
        const uint64_t pc = p < total ? p : total - 1;  // a place inside the last segment
        // Every lane shuffles: a lane reading from an inactive lane gets nothing.
        const uint64_t ai = __shfl(ashift, (int)js[u], kWave) + pc;
        const uint64_t os = __shfl(oshift, (int)js[u], kWave);
        key[u] = A.ak[ai] + __shfl(kshift, (int)js[u], kWave);
        val[u] = __shfl(b, (int)js[u], kWave) * A.av[ai];
        dst[u] = p < total ? os + p : ~0ull;
      }
#pragma unroll
      for (uint32_t u = 0; u < kSingleItems; ++u) {
        if (dst[u] < A.capacity) {
          A.ck[dst[u]] = key[u];
          A.cv[dst[u]] = val[u];
        }
      }
    }
  }
}

// Per-instance output offsets and products; overflow check.
__global__ void batch_offsets_kernel(const uint64_t *__restrict__ b_off, uint32_t n_inst, uint64_t nb,
                                     const uint64_t *__restrict__ seg_idx, const uint64_t *__restrict__ out_off,
                                     const uint64_t *__restrict__ prod_off, uint64_t capacity,
                                     uint64_t *__restrict__ ctr, uint64_t *__restrict__ c_off,
                                     uint64_t *__restrict__ nprod) {
  const uint64_t nc = ctr[kCtrNc];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n_inst;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t p = b_off[i];
    c_off[i] = (i < n_inst && p < nb) ? out_off[seg_idx[p]] : nc;
    if (i < n_inst) nprod[i] = prod_off[b_off[i + 1]] - prod_off[p];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && nc > capacity) ctr[kCtrOverflow] = 1;
}

__global__ void batch_clear_kernel(uint64_t *__restrict__ ctr) {
  if (threadIdx.x < kCtrWords) ctr[threadIdx.x] = 0;
}

__global__ void batch_zero_u32_kernel(uint32_t *__restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

inline uint32_t grid_for(uint64_t n, uint32_t cap = 16384) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  return (uint32_t)(b > cap ? cap : b);
}

struct Carve {
  char *p;
  template <typename T>
  T *take(uint64_t count) {
    T *r = (T *)p;
    p += ((count * sizeof(T) + 255) / 256) * 256;
    return r;
  }
};

}  // namespace

// ------------------------------------------------------------------- host
struct BatchTuning {
  uint32_t item_products = 1024;
  uint32_t small_max = 1024;
  uint32_t window = 1280;
  uint32_t blocks_per_cu = 3;
  int64_t piece_slots = -1;  // at most this many row pieces (-1: the whole pool)
  uint32_t dense_min = 256;  // smallest segment of a dense instance taken densely
};

}  // namespace lnt

struct lnt_batch_s {
  lnt_context ctx = nullptr;
  uint32_t count = 0;
  uint64_t na = 0, nb = 0;
  uint64_t a_bound = 0, b_bound = 0;
  uint64_t max_h = 0;  // largest A lattice height (row count) of the instances
  int cus = 256;
  lnt::BatchTuning tuning;
  lnt::DeviceBuffer meta;  // InstDev[count], RelinDev[2*count], a_off/b_off[count+1]
  lnt::DeviceBuffer ws;    // per-entry tables
  lnt::DeviceBuffer scan_tmp;
  lnt::DeviceBuffer ctrl;  // counters, c_off[count+1], nprod[count]
  // Row-piece slots (kPieceSlotBytes each): the pool of eager calls, and the
  // recorded graph's own (grown by the sizing run of a capture only).
  lnt::DeviceBuffer pieces, g_pieces;
  uint64_t piece_cap = 0, g_piece_cap = 0;
  uint64_t piece_need = 0, g_piece_need = 0;  // slots the last run asked for
  uint64_t *pinned = nullptr;  // host mirror of counters + c_off + nprod
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t own_stream = nullptr;
  int64_t graph_capacity = 0;
  bool symbolic_ready = false;  // tables of the last symbolic phase are live
  // Sort scratch of the recorded graph (ping-pong buffers, look-back status
  // words and their epoch, histograms, counters and flags, merge splits): the
  // batch's own, so that later calls on the context, which grow, trim or
  // reuse the context's scratch, never touch memory a replay writes.
  lnt::DeviceBuffer g_alt_keys, g_alt_vals, g_alt2_keys, g_alt2_vals, g_status, g_small, g_hist_part,
      g_relin_dev, g_aux;
  uint32_t g_epoch = 0;
};

namespace lnt {
namespace {

struct BatchPtrs {
  InstDev *inst;
  uint32_t *inst_dense;  // per instance: 1 = dense accumulator
  RelinDev *relin_a, *relin_b;
  uint64_t *a_off, *b_off;
  uint64_t *ka, *kb;
  double *va, *vb;
  uint64_t *a_lo, *a_base, *colkey, *prod_off, *seg_idx, *seg_first, *seg_u, *out_off, *small_rank, *sp_off, *cur,
      *large_rank;
  uint32_t *a_len, *einst, *head, *kind, *is_small, *is_large, *small_list, *np_list, *list3, *wcnt, *woff;
  uint64_t *dense_flag, *dense_rank;
  DenseDesc *dense_desc;
  uint64_t *ctr, *c_off, *nprod;
  PieceArgs pieces;
};

size_t ws_bytes(uint64_t na, uint64_t nb) {
  const uint64_t n1 = nb + 1;
  return 256 * 56 + 16 * na + n1 * (8 + 8 + 8 * 17 + 4 * 13 + sizeof(DenseDesc));
}

void carve(lnt_batch b, BatchPtrs *P) {
  Carve m{(char *)b->meta.ptr};
  P->inst = m.take<InstDev>(b->count);
  P->inst_dense = m.take<uint32_t>(b->count);
  P->relin_a = m.take<RelinDev>(b->count);
  P->relin_b = m.take<RelinDev>(b->count);
  P->a_off = m.take<uint64_t>(b->count + 1);
  P->b_off = m.take<uint64_t>(b->count + 1);
  Carve c{(char *)b->ws.ptr};
  const uint64_t n1 = b->nb + 1;
  P->ka = c.take<uint64_t>(b->na);
  P->va = c.take<double>(b->na);
  P->kb = c.take<uint64_t>(b->nb);
  P->vb = c.take<double>(b->nb);
  P->a_lo = c.take<uint64_t>(n1);
  P->a_base = c.take<uint64_t>(n1);
  P->colkey = c.take<uint64_t>(n1);
  P->prod_off = c.take<uint64_t>(n1);
  P->seg_idx = c.take<uint64_t>(n1);
  P->seg_first = c.take<uint64_t>(n1);
  P->seg_u = c.take<uint64_t>(n1);
  P->out_off = c.take<uint64_t>(n1);
  P->small_rank = c.take<uint64_t>(n1);
  P->sp_off = c.take<uint64_t>(n1);
  P->cur = c.take<uint64_t>(n1);
  P->a_len = c.take<uint32_t>(n1);
  P->einst = c.take<uint32_t>(n1);
  P->head = c.take<uint32_t>(n1);
  P->kind = c.take<uint32_t>(n1);
  P->is_small = c.take<uint32_t>(n1);
  P->small_list = c.take<uint32_t>(n1);
  P->np_list = c.take<uint32_t>(n1);
  P->list3 = c.take<uint32_t>(n1);
  P->wcnt = c.take<uint32_t>(n1);
  P->woff = c.take<uint32_t>(n1);
  P->large_rank = c.take<uint64_t>(n1);
  P->is_large = c.take<uint32_t>(n1);
  P->dense_flag = c.take<uint64_t>(n1);
  P->dense_rank = c.take<uint64_t>(n1);
  P->dense_desc = c.take<DenseDesc>(n1);
  P->pieces.rmin = c.take<uint64_t>(n1);
  P->pieces.width = c.take<uint64_t>(n1);
  P->pieces.pbase = c.take<uint64_t>(n1);
  P->pieces.pq = c.take<uint32_t>(n1);
  P->pieces.list3w = c.take<uint32_t>(n1);
  Carve q{(char *)b->pieces.ptr};
  P->pieces.piece_seg = q.take<uint32_t>(b->piece_cap);
  P->pieces.piece_u = q.take<uint32_t>(b->piece_cap);
  P->pieces.piece_ent = q.take<uint64_t>(b->piece_cap);
  P->pieces.status = q.take<uint64_t>(b->piece_cap);
  P->pieces.bnd = q.take<uint32_t>(b->piece_cap * kPieceBnd);
  P->pieces.cap = b->tuning.piece_slots < 0 ? b->piece_cap
                                             : std::min<uint64_t>(b->piece_cap, (uint64_t)b->tuning.piece_slots);
  Carve k{(char *)b->ctrl.ptr};
  P->ctr = k.take<uint64_t>(kCtrWords);
  P->c_off = k.take<uint64_t>(b->count + 1);
  P->nprod = k.take<uint64_t>(b->count);
}

// Everything up to and including the column tables and the symbolic pass
// (when `symbolic`); no host synchronisation.
int enqueue_front(lnt_batch b, const BatchPtrs &P, const int64_t *a_keys, const double *a_vals,
                  const int64_t *b_keys, const double *b_vals, bool symbolic, bool direct = false) {
  lnt_context ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const uint64_t nb = b->nb;
  batch_clear_kernel<<<1, 64, 0, st>>>(P.ctr);
  LNT_HIP(hipGetLastError());
  if (b->na)
    batch_relin_kernel<<<grid_for(b->na), kBlock, 0, st>>>(a_keys, a_vals, b->na, P.a_off, b->count, P.inst, 0,
                                                           P.relin_a, P.ka, P.va, P.ctr);
  if (nb)
    batch_relin_kernel<<<grid_for(nb), kBlock, 0, st>>>(b_keys, b_vals, nb, P.b_off, b->count, P.inst, 1,
                                                        P.relin_b, P.kb, P.vb, P.ctr);
  LNT_HIP(hipGetLastError());
  if (b->na > 1) LNT_TRY(sort_device(ctx, P.ka, P.va, (uint32_t)b->na, b->a_bound, nullptr));
  if (nb > 1) LNT_TRY(sort_device(ctx, P.kb, P.vb, (uint32_t)nb, b->b_bound, nullptr));
  if (nb == 0) return LNT_OK;
  batch_entry_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.ka, P.kb, nb, P.b_off, b->count, P.inst, P.a_lo, P.a_len,
                                                      P.a_base, P.colkey, P.einst, P.head);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, P.a_len, P.prod_off, nb, P.prod_off + nb, b->scan_tmp, 0));
  LNT_TRY(exclusive_scan<uint32_t>(ctx, P.head, P.seg_idx, nb, P.ctr + kCtrNseg, b->scan_tmp, 0));
  if (!symbolic) return LNT_OK;
  batch_seg_first_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.head, P.seg_idx, nb, P.seg_first);
  const uint32_t small_max = std::min<uint32_t>(b->tuning.small_max, kCap / 2);
  batch_classify_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.seg_first, P.prod_off, P.einst, P.inst, P.inst_dense,
                                                         nb, small_max, b->tuning.dense_min, P.ctr, P.kind,
                                                         P.is_small, P.seg_u, P.is_large, P.dense_flag);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint64_t>(ctx, P.dense_flag, P.dense_rank, nb, P.ctr + kCtrNdense, b->scan_tmp, 0));
  batch_dense_list_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.dense_flag, P.dense_rank, P.seg_first, P.prod_off,
                                                           P.einst, P.inst, P.ctr, nb, P.dense_desc);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, P.is_large, P.large_rank, nb, P.ctr + kCtrN3, b->scan_tmp, 0));
  batch_large_list_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.is_large, P.large_rank, nb, P.list3);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, P.is_small, P.small_rank, nb, P.ctr + kCtrNsmall, b->scan_tmp, 0));
  batch_zero_u32_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.np_list, nb);
  batch_small_list_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.is_small, P.small_rank, P.seg_first, P.prod_off, nb,
                                                           P.ctr, P.small_list, P.np_list);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, P.np_list, P.sp_off, nb, P.ctr + kCtrSmallProd, b->scan_tmp, 0));
  ColumnArgs A{P.ka, P.va, P.vb, P.a_lo, P.a_base, P.colkey, P.prod_off, P.seg_first, P.out_off,
               P.a_len, P.einst, P.inst, nb, P.seg_u, nullptr, nullptr, 0};
  const uint32_t grid = (uint32_t)b->cus * b->tuning.blocks_per_cu;
  const uint32_t G = std::min<uint32_t>(b->tuning.item_products, kCap / 2);
  const uint32_t window = std::min<uint32_t>(b->tuning.window, 3 * kCap / 4);
  const uint32_t piece_grid = (uint32_t)b->cus * 3;
  batch_piece_plan_kernel<<<grid_for(nb * kWave), kBlock, 0, st>>>(A, P.list3, P.pieces, P.ctr);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, P.pieces.pq, P.pieces.pbase, nb, P.ctr + kCtrPieces, b->scan_tmp, 0));
  batch_piece_assign_kernel<<<grid_for(nb * kWave), kBlock, 0, st>>>(A, P.list3, P.pieces, P.ctr);
  batch_piece_bounds_kernel<<<grid_for(P.pieces.cap * kPieceBnd), kBlock, 0, st>>>(A, P.pieces, P.ctr);
  {
    // 32-bit hash keys (half the LDS: twice the workgroups per CU) when the
    // item's segment (at most kMaxSegsPerItem - 1 = 1024, so the key stays
    // below 0xFFFFFFFF) and the rows fit in 32 bits.
    const uint32_t row_bits = b->max_h > 1 ? 64u - (uint32_t)__builtin_clzll(b->max_h - 1) : 0u;
    if (row_bits + 11 <= 32)
      batch_small_kernel<false, uint32_t><<<std::max<uint32_t>(grid, (uint32_t)b->cus * 6), kBlock, 0, st>>>(
          A, P.small_list, P.sp_off, G, row_bits, P.ctr);
    else
      batch_small_kernel<false, uint64_t><<<grid, kBlock, 0, st>>>(A, P.small_list, P.sp_off, G, kRowBits, P.ctr);
  }
  batch_dense_kernel<false, 256, 8, kDenseH><<<(uint32_t)b->cus * 8, 256, 0, st>>>(A, P.dense_desc, 0, P.ctr);
  if (!direct) batch_piece_kernel<false, false><<<piece_grid, kPieceBlock, 0, st>>>(A, P.list3, P.pieces, P.ctr);
  batch_large_kernel<false, uint64_t><<<grid, kBlock, 0, st>>>(A, P.pieces.list3w, P.cur, P.wcnt, P.woff, window,
                                                               P.ctr, kCtrN3w);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint64_t>(ctx, P.seg_u, P.out_off, nb, P.ctr + kCtrNc, b->scan_tmp, 0));
  return LNT_OK;
}

int enqueue_numeric(lnt_batch b, const BatchPtrs &P, uint64_t *ck, double *cv, uint64_t capacity,
                    bool direct = false) {
  lnt_context ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const uint64_t nb = b->nb;
  if (nb) {
    ColumnArgs A{P.ka, P.va, P.vb, P.a_lo, P.a_base, P.colkey, P.prod_off, P.seg_first, P.out_off,
                 P.a_len, P.einst, P.inst, nb, P.seg_u, ck, cv, capacity};
    if (direct) {
      // The pieces first, at the other segments' offsets (out_off counts
      // every segment but the pieced ones), then every segment's offset.
      if (P.pieces.cap)
        batch_zero_u32_kernel<<<grid_for(2 * P.pieces.cap), kBlock, 0, st>>>((uint32_t *)P.pieces.status,
                                                                             2 * P.pieces.cap);
      batch_piece_kernel<true, true><<<(uint32_t)b->cus * 3, kPieceBlock, 0, st>>>(A, P.list3, P.pieces, P.ctr);
      batch_piece_total_kernel<<<grid_for(nb), kBlock, 0, st>>>(P.list3, P.pieces, P.seg_u, P.ctr);
      LNT_HIP(hipGetLastError());
      LNT_TRY(exclusive_scan<uint64_t>(ctx, P.seg_u, P.out_off, nb, P.ctr + kCtrNc, b->scan_tmp, 0));
    }
    const uint32_t grid = (uint32_t)b->cus * b->tuning.blocks_per_cu;
    const uint32_t G = std::min<uint32_t>(b->tuning.item_products, kCap / 2);
    const uint32_t window = std::min<uint32_t>(b->tuning.window, 3 * kCap / 4);
    batch_single_kernel<<<grid_for(nb) / 8 * 8 + 8, kBlock, 0, st>>>(A, P.kind, P.ctr);
    // Items hold at most kMaxSegsPerItem segments: 11 bits above the rows.
    const uint32_t row_bits = b->max_h > 1 ? 64u - (uint32_t)__builtin_clzll(b->max_h - 1) : 0u;
    if (row_bits + 11 <= 32)
      batch_small_kernel<true, uint32_t><<<grid, kBlock, 0, st>>>(A, P.small_list, P.sp_off, G, row_bits, P.ctr);
    else
      batch_small_kernel<true, uint64_t><<<grid, kBlock, 0, st>>>(A, P.small_list, P.sp_off, G, kRowBits, P.ctr);
    batch_dense_kernel<true, 256, 8, kDenseSmallH><<<(uint32_t)b->cus * 7, 256, 0, st>>>(A, P.dense_desc, 1, P.ctr);
    batch_dense_kernel<true, 512, 4, kDenseH><<<(uint32_t)b->cus * 2, 512, 0, st>>>(A, P.dense_desc, 2, P.ctr);
    if (!direct)
      batch_piece_kernel<true, false><<<(uint32_t)b->cus * 3, kPieceBlock, 0, st>>>(A, P.list3, P.pieces, P.ctr);
    if (b->max_h <= (uint64_t(1) << 32))
      batch_large_kernel<true, uint32_t><<<grid, kBlock, 0, st>>>(A, P.pieces.list3w, P.cur, P.wcnt, P.woff, window,
                                                                  P.ctr, kCtrN3w);
    else
      batch_large_kernel<true, uint64_t><<<grid, kBlock, 0, st>>>(A, P.pieces.list3w, P.cur, P.wcnt, P.woff, window,
                                                                  P.ctr, kCtrN3w);
    LNT_HIP(hipGetLastError());
  }
  batch_offsets_kernel<<<grid_for(b->count + 1), kBlock, 0, st>>>(P.b_off, b->count, nb, P.seg_idx, P.out_off,
                                                                  P.prod_off, capacity, P.ctr, P.c_off, P.nprod);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

// Tickets of the numeric pass start from zero again.
__global__ void batch_reset_tickets_kernel(uint64_t *__restrict__ ctr) {
  if (threadIdx.x == 0) {
    ctr[kCtrTicketLarge] = 0;
    ctr[kCtrTicketPiece] = 0;
    ctr[kCtrTicketSmall] = 0;
    ctr[kCtrTicketDense] = 0;
    ctr[kCtrTicketDense2] = 0;
  }
}

// The whole pipeline into arrays known in advance (a recorded graph): the
// row pieces write C directly (one pass instead of a counting and a writing
// one).
int enqueue_all(lnt_batch b, const BatchPtrs &P, const int64_t *a_keys, const double *a_vals, const int64_t *b_keys,
                const double *b_vals, uint64_t *ck, double *cv, uint64_t capacity) {
  LNT_TRY(enqueue_front(b, P, a_keys, a_vals, b_keys, b_vals, true, true));
  batch_reset_tickets_kernel<<<1, 64, 0, b->ctx->stream>>>(P.ctr);
  LNT_HIP(hipGetLastError());
  return enqueue_numeric(b, P, ck, cv, capacity, true);
}

// Copies counters, offsets and products to the host mirror and waits.
int read_back(lnt_batch b) {
  LNT_HIP(hipMemcpyAsync(b->pinned, b->ctrl.ptr, b->ctrl.bytes, hipMemcpyDeviceToHost, b->ctx->stream));
  LNT_HIP(hipStreamSynchronize(b->ctx->stream));
  return LNT_OK;
}

// c_off sits right after the counters in the control block (256-byte carve).
const uint64_t *host_c_off(lnt_batch b) { return b->pinned + ((kCtrWords * 8 + 255) / 256) * 256 / 8; }
const uint64_t *host_nprod(lnt_batch b) {
  return host_c_off(b) + (((b->count + 1) * 8 + 255) / 256) * 256 / 8;
}

int check_key_error(lnt_batch b) {
  if (b->pinned[kCtrKeyErr])
    return fail(LNT_ERR_KEY_RANGE, "a key is negative or not below its instance's dimensionality");
  return LNT_OK;
}

// Piece slots for a first run: one per 16 B entries (a large segment has at
// least two entries and usually more than one piece); a run that needs more
// reports its need, and the pool grows for the next run (pieces beyond the
// pool go to the windows meanwhile).
int ensure_pieces(lnt_batch b, uint64_t need) {
  if (b->piece_cap == 0) b->piece_cap = std::max<uint64_t>(65536, b->nb / 16 + 1);
  if (need > b->piece_cap) b->piece_cap = need + need / 8 + 1024;
  return ensure(b->pieces, (size_t)b->piece_cap * kPieceSlotBytes + 2048);
}

int prepare(lnt_batch b) {
  LNT_TRY(ensure_pieces(b, b->piece_need));
  LNT_TRY(ensure(b->ws, ws_bytes(b->na, b->nb)));
  LNT_TRY(ensure(b->scan_tmp, (scan_workspace_words(b->nb + 1) + 64) * 8));
  return LNT_OK;
}

// While alive, the batch's piece pool is the recorded graph's.
struct PieceSwap {
  lnt_batch b;
  explicit PieceSwap(lnt_batch bb) : b(bb) { swap_all(); }
  ~PieceSwap() { swap_all(); }
  void swap_all() {
    std::swap(b->pieces, b->g_pieces);
    std::swap(b->piece_cap, b->g_piece_cap);
    std::swap(b->piece_need, b->g_piece_need);
  }
};

struct StreamSwap {
  lnt_context ctx;
  hipStream_t saved;
  StreamSwap(lnt_context c, hipStream_t s) : ctx(c), saved(c->stream) { ctx->stream = s; }
  ~StreamSwap() { ctx->stream = saved; }
};

// While alive, the context's sort scratch is the batch's graph scratch (and
// the other way round): the sizing run and the capture size and record the
// batch's buffers, and the context gets its own back unchanged.
struct ScratchSwap {
  lnt_context ctx;
  lnt_batch b;
  ScratchSwap(lnt_context c, lnt_batch bb) : ctx(c), b(bb) { swap_all(); }
  ~ScratchSwap() { swap_all(); }
  void swap_all() {
    std::swap(ctx->alt_keys, b->g_alt_keys);
    std::swap(ctx->alt_vals, b->g_alt_vals);
    std::swap(ctx->alt2_keys, b->g_alt2_keys);
    std::swap(ctx->alt2_vals, b->g_alt2_vals);
    std::swap(ctx->status, b->g_status);
    std::swap(ctx->small, b->g_small);
    std::swap(ctx->hist_part, b->g_hist_part);
    std::swap(ctx->relin_dev, b->g_relin_dev);
    std::swap(ctx->aux, b->g_aux);
    std::swap(ctx->epoch, b->g_epoch);
  }
};

}  // namespace
}  // namespace lnt

using namespace lnt;

extern "C" {

int lnt_batch_create(lnt_context ctx, const lnt_contraction *items, int64_t count, lnt_batch *out) {
  if (!ctx || !out || (count > 0 && !items)) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (count < 0 || count >= ((int64_t)1 << 31)) return fail(LNT_ERR_INVALID_ARGUMENT, "count out of range");
  LNT_HIP(hipSetDevice(ctx->device));
  std::vector<InstDev> inst(count);
  std::vector<RelinDev> ra(count), rb(count);
  std::vector<uint64_t> a_off(count + 1), b_off(count + 1);
  unsigned __int128 abound = 0, bbound = 0;
  uint64_t na = 0, nb = 0, max_h = 0;
  for (int64_t i = 0; i < count; ++i) {
    const lnt_contraction &c = items[i];
    const std::string at = "instance " + std::to_string(i) + ": ";
    Relin r1{}, r2{};
    uint64_t da = 0, db = 0;
    int rc = make_relin(c.order_a, c.dims_a, c.cur_seq_a, c.new_seq_a, &r1, &da);
    if (rc != LNT_OK) return fail(rc, at + lnt_last_error());
    rc = make_relin(c.order_b, c.dims_b, c.cur_seq_b, c.new_seq_b, &r2, &db);
    if (rc != LNT_OK) return fail(rc, at + lnt_last_error());
    for (int q = 0; q < 3; ++q)
      if (c.lat_a[q] < 1 || c.lat_b[q] < 1) return fail(LNT_ERR_INVALID_ARGUMENT, at + "lattice dims must be >= 1");
    if (c.lat_a[1] != c.lat_b[0] || c.lat_a[2] != c.lat_b[2])
      return fail(LNT_ERR_DIMENSION_MISMATCH, at + "A's width and depth must equal B's height and depth");
    unsigned __int128 pa = (unsigned __int128)c.lat_a[0] * c.lat_a[1] * c.lat_a[2];
    unsigned __int128 pb = (unsigned __int128)c.lat_b[0] * c.lat_b[1] * c.lat_b[2];
    unsigned __int128 pc = (unsigned __int128)c.lat_a[0] * c.lat_b[1] * c.lat_b[2];
    if (pa != da || pb != db)
      return fail(LNT_ERR_DIMENSION_MISMATCH, at + "lattice shape does not match the operand's dimensionality");
    if (pc > ((unsigned __int128)1 << 63)) return fail(LNT_ERR_TOO_LARGE, at + "C's dimensionality exceeds 2^63");
    if (c.nnz_a < 0 || c.nnz_b < 0 || (uint64_t)c.nnz_a > da || (uint64_t)c.nnz_b > db)
      return fail(LNT_ERR_INVALID_ARGUMENT, at + "nnz must lie in [0, dimensionality]");
    InstDev &I = inst[i];
    I.a_off = na;
    I.na = (uint64_t)c.nnz_a;
    I.b_off = nb;
    I.nb = (uint64_t)c.nnz_b;
    I.a_base = (uint64_t)abound;
    I.b_base = (uint64_t)bbound;
    I.a_dim = da;
    I.b_dim = db;
    I.h = (uint64_t)c.lat_a[0];
    max_h = std::max(max_h, I.h);
    I.w = (uint64_t)c.lat_a[1];
    I.bw = (uint64_t)c.lat_b[1];
    I.div_w = make_divider(I.w);
    I.div_wbw = make_divider(I.w * I.bw);
    a_off[i] = na;
    b_off[i] = nb;
    na += I.na;
    nb += I.nb;
    abound += da;
    bbound += db;
    if (abound >= ((unsigned __int128)1 << 63) || bbound >= ((unsigned __int128)1 << 63))
      return fail(LNT_ERR_TOO_LARGE, "the instances' dimensionalities sum past 2^63");
    if (na >= (uint64_t(1) << 32) || nb >= (uint64_t(1) << 32))
      return fail(LNT_ERR_TOO_LARGE, "a batch holds fewer than 2^32 entries per operand");
    ra[i] = r1.active ? relin_for_device(r1) : RelinDev{};
    rb[i] = r2.active ? relin_for_device(r2) : RelinDev{};
    if (!r1.active) ra[i].weight[0] = 1;  // order 0: the key times 1
    if (!r2.active) rb[i].weight[0] = 1;
  }
  a_off[count] = na;
  b_off[count] = nb;

  lnt_batch b = new lnt_batch_s();
  b->ctx = ctx;
  b->count = (uint32_t)count;
  b->na = na;
  b->max_h = max_h;
  b->nb = nb;
  b->a_bound = std::max<uint64_t>((uint64_t)abound, 1);
  b->b_bound = std::max<uint64_t>((uint64_t)bbound, 1);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && cus > 0)
    b->cus = cus;
  auto bail = [&](int rc) {
    lnt_batch_destroy(b);
    return rc;
  };
  const size_t meta = (((count * sizeof(InstDev) + 255) / 256) + ((count * 4 + 255) / 256) +
                       2 * ((count * sizeof(RelinDev) + 255) / 256) + 2 * (((count + 1) * 8 + 255) / 256)) * 256 +
                       256;
  int rc = ensure(b->meta, meta);
  if (rc != LNT_OK) return bail(rc);
  const size_t ctrl = (((kCtrWords * 8 + 255) / 256) + (((count + 1) * 8 + 255) / 256) + ((count * 8 + 255) / 256)) * 256 + 256;
  rc = ensure(b->ctrl, ctrl);
  if (rc != LNT_OK) return bail(rc);
  if (hipHostMalloc((void **)&b->pinned, ctrl) != hipSuccess) {
    (void)hipGetLastError();
    b->pinned = nullptr;
    return bail(fail(LNT_ERR_OUT_OF_MEMORY, "hipHostMalloc of the batch's control block failed"));
  }
  BatchPtrs P{};
  carve(b, &P);
  auto up = [&](void *dst, const void *src, size_t bytes) -> int {
    if (!bytes) return LNT_OK;
    LNT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return LNT_OK;
  };
  if ((rc = up(P.inst, inst.data(), count * sizeof(InstDev))) != LNT_OK) return bail(rc);
  {
    const std::vector<uint32_t> none(count, 0u);  // every instance on the batch's own paths
    if ((rc = up(P.inst_dense, none.data(), count * 4)) != LNT_OK) return bail(rc);
  }
  if ((rc = up(P.relin_a, ra.data(), count * sizeof(RelinDev))) != LNT_OK) return bail(rc);
  if ((rc = up(P.relin_b, rb.data(), count * sizeof(RelinDev))) != LNT_OK) return bail(rc);
  if ((rc = up(P.a_off, a_off.data(), (count + 1) * 8)) != LNT_OK) return bail(rc);
  if ((rc = up(P.b_off, b_off.data(), (count + 1) * 8)) != LNT_OK) return bail(rc);
  *out = b;
  return LNT_OK;
}

int lnt_batch_destroy(lnt_batch b) {
  if (!b) return LNT_OK;
  if (b->ctx) (void)hipSetDevice(b->ctx->device);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  release(b->meta);
  release(b->ws);
  release(b->scan_tmp);
  release(b->ctrl);
  release(b->pieces);
  release(b->g_pieces);
  for (DeviceBuffer *x : {&b->g_alt_keys, &b->g_alt_vals, &b->g_alt2_keys, &b->g_alt2_vals, &b->g_status,
                          &b->g_small, &b->g_hist_part, &b->g_relin_dev, &b->g_aux})
    release(*x);
  if (b->pinned) (void)hipHostFree(b->pinned);
  delete b;
  return LNT_OK;
}

int lnt_batch_symbolic(lnt_batch b, const int64_t *a_keys, const double *a_vals, const int64_t *b_keys,
                       const double *b_vals, int64_t *nnz_c) {
  if (!b || !nnz_c) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  if ((b->na && (!a_keys || !a_vals)) || (b->nb && (!b_keys || !b_vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null operand");
  b->symbolic_ready = false;
  LNT_HIP(hipSetDevice(b->ctx->device));
  LNT_TRY(prepare(b));
  BatchPtrs P{};
  carve(b, &P);
  LNT_TRY(enqueue_front(b, P, a_keys, a_vals, b_keys, b_vals, true));
  LNT_TRY(read_back(b));
  LNT_TRY(check_key_error(b));
  *nnz_c = (int64_t)b->pinned[kCtrNc];
  b->symbolic_ready = true;
  // The numeric phase uses the pool as it is; a larger need grows it after.
  b->piece_need = b->pinned[kCtrPieces];
  return LNT_OK;
}

int lnt_batch_numeric(lnt_batch b, int64_t *c_keys, double *c_vals, int64_t capacity, int64_t *c_offsets) {
  if (!b || capacity < 0 || (capacity > 0 && (!c_keys || !c_vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null argument or negative capacity");
  if (!b->symbolic_ready) return fail(LNT_ERR_INVALID_ARGUMENT, "lnt_batch_numeric needs lnt_batch_symbolic first");
  const uint64_t nc = b->pinned[kCtrNc];
  if ((uint64_t)capacity < nc)
    return fail(LNT_ERR_INVALID_ARGUMENT, "capacity " + std::to_string(capacity) + " is below nnz(C) = " +
                                              std::to_string(nc));
  b->symbolic_ready = false;
  LNT_HIP(hipSetDevice(b->ctx->device));
  BatchPtrs P{};
  carve(b, &P);
  batch_reset_tickets_kernel<<<1, 64, 0, b->ctx->stream>>>(P.ctr);
  LNT_HIP(hipGetLastError());
  LNT_TRY(enqueue_numeric(b, P, (uint64_t *)c_keys, c_vals, (uint64_t)capacity));
  LNT_TRY(read_back(b));
  if (c_offsets) {
    const uint64_t *co = host_c_off(b);
    for (uint32_t i = 0; i <= b->count; ++i) c_offsets[i] = (int64_t)co[i];
  }
  return LNT_OK;
}

int lnt_batch_contract(lnt_batch b, const int64_t *a_keys, const double *a_vals, const int64_t *b_keys,
                       const double *b_vals, int64_t **c_keys, double **c_vals, int64_t *c_offsets) {
  if (!b || !c_keys || !c_vals || !c_offsets) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  *c_keys = nullptr;
  *c_vals = nullptr;
  int64_t nc = 0;
  LNT_TRY(lnt_batch_symbolic(b, a_keys, a_vals, b_keys, b_vals, &nc));
  void *k = nullptr, *v = nullptr;
  const size_t bytes = (nc ? (size_t)nc : 1) * 8;
  if (hipMalloc(&k, bytes) != hipSuccess || hipMalloc(&v, bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (k) (void)hipFree(k);
    b->symbolic_ready = false;
    return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of the batch result (" + std::to_string(nc) + " entries) failed");
  }
  int rc = lnt_batch_numeric(b, (int64_t *)k, (double *)v, nc, c_offsets);
  if (rc != LNT_OK) {
    (void)hipFree(k);
    (void)hipFree(v);
    return rc;
  }
  *c_keys = (int64_t *)k;
  *c_vals = (double *)v;
  return LNT_OK;
}

int lnt_batch_nprod(lnt_batch b, const int64_t *a_keys, const int64_t *b_keys, int64_t *products) {
  if (!b || !products) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  if ((b->na && !a_keys) || (b->nb && !b_keys)) return fail(LNT_ERR_INVALID_ARGUMENT, "null operand");
  b->symbolic_ready = false;
  LNT_HIP(hipSetDevice(b->ctx->device));
  if (b->nb == 0) {
    for (uint32_t i = 0; i < b->count; ++i) products[i] = 0;
    return LNT_OK;
  }
  LNT_TRY(prepare(b));
  BatchPtrs P{};
  carve(b, &P);
  LNT_TRY(enqueue_front(b, P, a_keys, nullptr, b_keys, nullptr, false));
  batch_offsets_kernel<<<grid_for(b->count + 1), kBlock, 0, b->ctx->stream>>>(
      P.b_off, b->count, b->nb, P.seg_idx, P.prod_off, P.prod_off, ~0ull, P.ctr, P.c_off, P.nprod);
  LNT_HIP(hipGetLastError());
  LNT_TRY(read_back(b));
  LNT_TRY(check_key_error(b));
  const uint64_t *np = host_nprod(b);
  for (uint32_t i = 0; i < b->count; ++i) products[i] = (int64_t)np[i];
  return LNT_OK;
}

int lnt_batch_capture(lnt_batch b, const int64_t *a_keys, const double *a_vals, const int64_t *b_keys,
                      const double *b_vals, int64_t *c_keys, double *c_vals, int64_t capacity, int64_t *nnz_c) {
  if (!b || (capacity > 0 && (!c_keys || !c_vals)) || capacity < 0)
    return fail(LNT_ERR_INVALID_ARGUMENT, "null argument or negative capacity");
  if ((b->na && (!a_keys || !a_vals)) || (b->nb && (!b_keys || !b_vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "null operand");
  lnt_context ctx = b->ctx;
  b->symbolic_ready = false;
  LNT_HIP(hipSetDevice(ctx->device));
  // The old graph goes first: its scratch may be regrown below.
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  // The graph's piece pool is its own: eager calls may regrow theirs.
  PieceSwap piece_swap(b);
  LNT_TRY(prepare(b));
  BatchPtrs P{};
  carve(b, &P);
  // From here to the end of the capture the context sorts in the batch's own
  // scratch (its stream is drained first: queued work may still use the
  // context's buffers, which stay where they are).
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  ScratchSwap scratch(ctx, b);
  // One unrecorded run sizes every workspace (sort scratch, status words)
  // and checks the operands.
  LNT_TRY(enqueue_all(b, P, a_keys, a_vals, b_keys, b_vals, (uint64_t *)c_keys, c_vals, (uint64_t)capacity));
  LNT_TRY(read_back(b));
  LNT_TRY(check_key_error(b));
  if (nnz_c) *nnz_c = (int64_t)b->pinned[kCtrNc];
  // Room for every piece of these operands in the recorded graph.
  if (b->pinned[kCtrPieces] > b->piece_cap) {
    LNT_TRY(ensure_pieces(b, b->pinned[kCtrPieces]));
    carve(b, &P);
  }
  // The legacy null stream cannot be captured: record on a stream of our own
  // (blocking, default priority) when the context has none.
  hipStream_t st = ctx->stream;
  if (!st) {
    if (!b->own_stream) LNT_HIP(hipStreamCreate(&b->own_stream));
    st = b->own_stream;
  }
  int rc = LNT_OK;
  {
    StreamSwap swap(ctx, st);
    LNT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    ctx->capturing = 1;
    set_alloc_forbidden(true);
    rc = enqueue_all(b, P, a_keys, a_vals, b_keys, b_vals, (uint64_t *)c_keys, c_vals, (uint64_t)capacity);
    set_alloc_forbidden(false);
    ctx->capturing = 0;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc == LNT_OK && e != hipSuccess) rc = fail(LNT_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (rc == LNT_OK) {
      e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = fail(LNT_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    if (g && rc == LNT_OK) b->graph = g;
    else if (g) (void)hipGraphDestroy(g);
  }
  b->graph_capacity = capacity;
  return rc;
}

int lnt_batch_launch(lnt_batch b) {
  if (!b) return fail(LNT_ERR_INVALID_ARGUMENT, "null batch");
  b->symbolic_ready = false;
  if (!b->exec) return fail(LNT_ERR_INVALID_ARGUMENT, "no recorded graph: call lnt_batch_capture first");
  LNT_HIP(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream ? b->ctx->stream : b->own_stream;
  LNT_HIP(hipGraphLaunch(b->exec, st));
  return LNT_OK;
}

int lnt_batch_replay(lnt_batch b, int64_t *c_offsets) {
  LNT_TRY(lnt_batch_launch(b));
  hipStream_t st = b->ctx->stream ? b->ctx->stream : b->own_stream;
  LNT_HIP(hipMemcpyAsync(b->pinned, b->ctrl.ptr, b->ctrl.bytes, hipMemcpyDeviceToHost, st));
  LNT_HIP(hipStreamSynchronize(st));
  LNT_TRY(check_key_error(b));
  if (b->pinned[kCtrOverflow])
    return fail(LNT_ERR_TOO_LARGE, "the batch's result (" + std::to_string(b->pinned[kCtrNc]) +
                                       " entries) exceeds the recorded capacity " + std::to_string(b->graph_capacity));
  if (c_offsets) {
    const uint64_t *co = host_c_off(b);
    for (uint32_t i = 0; i <= b->count; ++i) c_offsets[i] = (int64_t)co[i];
  }
  return LNT_OK;
}

int lnt_batch_set_param(lnt_batch b, const char *name, int64_t value) {
  if (!b || !name) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  std::string n(name);
  auto in = [&](int64_t lo, int64_t hi) { return value >= lo && value <= hi; };
  if (n == "item_products" && in(64, kCap / 2)) b->tuning.item_products = (uint32_t)value;
  else if (n == "small_max" && in(2, kCap / 2)) b->tuning.small_max = (uint32_t)value;
  else if (n == "window" && in(64, 3 * kCap / 4)) b->tuning.window = (uint32_t)value;
  else if (n == "blocks_per_cu" && in(1, 8)) b->tuning.blocks_per_cu = (uint32_t)value;
  else if (n == "piece_slots" && value >= -1) b->tuning.piece_slots = value;
  else if (n == "dense_min" && in(16, (int64_t)1 << 30)) b->tuning.dense_min = (uint32_t)value;
  else return fail(LNT_ERR_INVALID_ARGUMENT, "unknown batch parameter or value out of range: " + n);
  return LNT_OK;
}

int lnt_batch_set_accumulators(lnt_batch b, const int32_t *acc) {
  if (!b || (b->count && !acc)) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<uint32_t> dense(b->count);
  for (uint32_t i = 0; i < b->count; ++i) {
    if (acc[i] < 0 || acc[i] >= LNT_NUM_ACCUMULATORS)
      return fail(LNT_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": unknown accumulator " +
                                                std::to_string(acc[i]));
    dense[i] = acc[i] == LNT_ACC_DENSE || acc[i] == LNT_ACC_MTT;
  }
  if (!b->count) return LNT_OK;
  LNT_HIP(hipSetDevice(b->ctx->device));
  BatchPtrs P{};
  carve(b, &P);
  // Stream ordered: calls already queued keep the choices they were made with.
  LNT_HIP(hipMemcpyAsync(P.inst_dense, dense.data(), (size_t)b->count * 4, hipMemcpyHostToDevice, b->ctx->stream));
  LNT_HIP(hipStreamSynchronize(b->ctx->stream));
  return LNT_OK;
}

int lnt_batch_get_param(lnt_batch b, const char *name, int64_t *value) {
  if (!b || !name || !value) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  std::string n(name);
  if (n == "item_products") *value = b->tuning.item_products;
  else if (n == "small_max") *value = b->tuning.small_max;
  else if (n == "window") *value = b->tuning.window;
  else if (n == "blocks_per_cu") *value = b->tuning.blocks_per_cu;
  else if (n == "piece_slots") *value = b->tuning.piece_slots;
  else if (n == "dense_min") *value = b->tuning.dense_min;
  else if (n == "dense_segments")
    *value = b->pinned ? (int64_t)((b->pinned[kCtrNdense] & 0xFFFFFFFFull) + (b->pinned[kCtrNdense] >> 32)) : 0;
  else return fail(LNT_ERR_INVALID_ARGUMENT, "unknown batch parameter: " + n);
  return LNT_OK;
}

}  // extern "C"
