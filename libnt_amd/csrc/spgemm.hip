// spgemm.hip -- sparse lattice product C = A * B per tab for gfx950.
//
// Replaces SparseLatticeBase::csc_no_accum<true> (multi_index_array/
// SparseLatticeBase.h:1854-1990, scatter :1993-2031): for every tab t and
// every column j of B, C(:,j,t) = sum_k B(k,j,t) * A(:,k,t), computed by
// expanding the products, ordering them by row, and summing equal rows.
// Lattice keys are column major, key = row + h*col + h*w*tab
// (SparseLatticeBase.h:867-877); A (h x w x d) and B (w x bw x d) arrive sorted
// ascending, C (h x bw x d) leaves sorted ascending.  Products are b*a as in
// mult_scatter; equal rows are summed in ascending (B row, A row) order.
//
// Pipeline (one output column of B = one "segment"):
//   1. per B entry: A column range by binary search, column id, head flag;
//   2. scans: product offsets per entry, segment ids;
//   3. per segment: classify by product count --
//        one B entry     -> C(:,j) is that A column scaled: no sort at all;
//        <= kSmall prods -> one workgroup expands into LDS, bitonic-sorts
//                           (row, position) and sums runs ("sort" strategy), or
//                           accumulates in an LDS hash table and sorts only the
//                           unique rows ("hash" strategy);
//        larger          -> all such segments together: products expanded to
//                           HBM with composite key (segment rank, row), one
//                           onesweep radix sort, run sums;
//   4. scan of per-segment output counts, then one coalesced write of C with
//      the lattice keys (row + h * (t*bw + j)).
// The other accumulators of the reference's strategy set: the dense sparse
// accumulator of csc_times (:2224-2351) and of mtt_times over compressed rows
// (:1579-1813) take every multi-entry segment in LDS when the index space
// fits (dense_kernel); outer_times's expand-sort-collect (:2034-2183) sends
// every multi-entry segment through the global sort.  A's columns are found by
// a binary search in its keys (DCSC) or through a column-pointer array (CSC,
// MultHelper<.,false>, :121-169).
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
constexpr int kSmall = 2048;  // largest segment handled inside one workgroup's LDS
constexpr uint32_t kSegGrid = 2048;  // workgroups of the short-segment kernels (they stride)

__device__ inline uint64_t lower_bound_u64(const uint64_t *a, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Largest i in [lo, hi) with off[i] <= x (off ascending, off[lo] <= x).
__device__ inline uint64_t upper_index(const uint64_t *off, uint64_t lo, uint64_t hi, uint64_t x) {
  while (hi - lo > 1) {
    uint64_t mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct LatticeDims {
  uint64_t ah, aw;          // A: h x w
  uint64_t bw;              // B: w x bw
  Divider b_tab;            // bh*bw
  Divider b_h;              // bh (= aw)
};

// A's column pointers (CSC): cp[c] = first A entry of global column c = col +
// w*tab, for every c in [0, w*d], by one search per column.
__global__ void col_ptr_kernel(const uint64_t *__restrict__ ak, uint64_t na, uint64_t ah, uint64_t ncols,
                               uint64_t *__restrict__ cp) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= ncols; c += (uint64_t)gridDim.x * blockDim.x)
    cp[c] = lower_bound_u64(ak, na, c * ah);
}

// 1. one thread per B entry.  cp: A's column pointers (CSC), or null to
// search A's keys for every entry (DCSC).
__global__ void entry_kernel(const uint64_t *__restrict__ ak, uint64_t na,
                             const uint64_t *__restrict__ bk, uint64_t nb, LatticeDims L,
                             const uint64_t *__restrict__ cp,
                             uint64_t *__restrict__ a_lo, uint32_t *__restrict__ a_len,
                             uint64_t *__restrict__ a_base, uint64_t *__restrict__ colid,
                             uint32_t *__restrict__ head) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < nb; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t key = bk[i];
    uint64_t t = divide(key, L.b_tab);
    uint64_t rem = key - t * L.b_tab.d;
    uint64_t col = divide(rem, L.b_h);
    uint64_t k = rem - col * L.b_h.d;
    uint64_t start = t * L.ah * L.aw + k * L.ah;
    uint64_t lo, hi;
    if (cp) {
      lo = cp[t * L.aw + k];
      hi = cp[t * L.aw + k + 1];
    } else {
      lo = lower_bound_u64(ak, na, start);
      hi = lower_bound_u64(ak, na, start + L.ah);
    }
    a_lo[i] = lo;
    a_len[i] = (uint32_t)(hi - lo);
    a_base[i] = start;
    uint64_t cid = t * L.bw + col;
    colid[i] = cid;
    uint32_t h = 1;
    if (i > 0) {
      uint64_t pk = bk[i - 1];
      uint64_t pt = divide(pk, L.b_tab);
      uint64_t pc = divide(pk - pt * L.b_tab.d, L.b_h);
      h = (pt * L.bw + pc) != cid;
    }
    head[i] = h;
  }
}

// 2b. segment starts: seg_first[seg_idx[i]] = i for head entries.
__global__ void seg_first_kernel(const uint32_t *__restrict__ head, const uint64_t *__restrict__ seg_idx,
                                 uint64_t nb, uint64_t *__restrict__ seg_first) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < nb; i += (uint64_t)gridDim.x * blockDim.x)
    if (head[i]) seg_first[seg_idx[i]] = i;
}

// 3. classification.  kind: 0 empty, 1 single B entry, 2 small, 3 large.
__global__ void classify_kernel(const uint64_t *__restrict__ seg_first, uint64_t nseg, uint64_t nb,
                                const uint64_t *__restrict__ prod_off, uint32_t small_max,
                                uint64_t *__restrict__ seg_u, uint32_t *__restrict__ list2,
                                uint32_t *__restrict__ large, uint8_t *__restrict__ direct,
                                uint32_t *__restrict__ counts, int mark_single) {
  uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; s < nseg; s += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t first = seg_first[s];
    uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
    uint64_t np = prod_off[last] - prod_off[first];
    uint32_t is_large = 0;
    // Marked segments are written by another kernel (single_copy_kernel, or
    // the pieces): the output kernel skips them.
    direct[s] = mark_single && last - first == 1 && np > 0;
    if (np == 0 || last - first == 1) {
      seg_u[s] = np;
    } else if (np <= small_max) {
      seg_u[s] = 0;
      list2[atomicAdd(&counts[0], 1u)] = (uint32_t)s;
    } else {
      seg_u[s] = 0;
      is_large = 1;
    }
    large[s] = is_large;
  }
}

// The large segments in ascending order (list3[pos[s]] = s): their pieces
// then run in output order.
__global__ void large_list_kernel(const uint32_t *__restrict__ large, const uint64_t *__restrict__ pos,
                                  uint64_t nseg, uint32_t *__restrict__ list3) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (uint64_t)gridDim.x * blockDim.x)
    if (large[s]) list3[pos[s]] = (uint32_t)s;
}

// Maps product p (0-based inside a segment) to its B entry e and A index.
__device__ inline void locate(const uint64_t *prod_off, uint64_t first, uint64_t last, uint64_t p,
                              const uint64_t *a_lo, uint64_t *e_out, uint64_t *a_idx) {
  uint64_t base = prod_off[first];
  uint64_t e = upper_index(prod_off, first, last, base + p);
  *e_out = e;
  *a_idx = a_lo[e] + (base + p - prod_off[e]);
}

// 3a. small segments, "sort" strategy: expand into LDS, bitonic sort of
// (row, position), sum runs in position order.
__global__ __launch_bounds__(kBlock) void small_sort_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, const double *__restrict__ bv,
    const uint64_t *__restrict__ a_lo, const uint64_t *__restrict__ a_base,
    const uint64_t *__restrict__ prod_off, const uint64_t *__restrict__ seg_first, uint64_t nseg,
    uint64_t nb, const uint32_t *__restrict__ list2, const uint32_t *__restrict__ n2_dev, uint64_t *__restrict__ seg_u,
    uint64_t *__restrict__ tmp_row, double *__restrict__ tmp_val) {
  __shared__ uint64_t s_row[kSmall];
  __shared__ uint16_t s_pos[kSmall];
  __shared__ double s_val[kSmall];
  __shared__ uint64_t s_scan[kBlock / kWave];
  const uint32_t tid = threadIdx.x;
  // Segments in a stride over the grid (their count is read on the device).
  const uint32_t n2 = *n2_dev;
  for (uint32_t it = blockIdx.x; it < n2; it += gridDim.x) {
  const uint64_t s = list2[it];
  const uint64_t first = seg_first[s];
  const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
  const uint64_t base = prod_off[first];
  const uint32_t np = (uint32_t)(prod_off[last] - base);
  uint32_t P = 1;
  while (P < np) P <<= 1;
  for (uint32_t p = tid; p < P; p += kBlock) {
    if (p < np) {
      uint64_t e, ai;
      locate(prod_off, first, last, p, a_lo, &e, &ai);
      s_row[p] = ak[ai] - a_base[e];
      s_val[p] = bv[e] * av[ai];
    } else {
      s_row[p] = ~0ull;
    }
    s_pos[p] = (uint16_t)p;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = tid; i < P; i += kBlock) {
        uint32_t l = i ^ j;
        if (l > i) {
          uint64_t ri = s_row[i], rl = s_row[l];
          uint16_t pi = s_pos[i], pl = s_pos[l];
          bool gt = ri > rl || (ri == rl && pi > pl);
          bool up = (i & k) == 0;
          if (gt == up) {
            s_row[i] = rl; s_row[l] = ri;
            s_pos[i] = pl; s_pos[l] = pi;
          }
        }
      }
      __syncthreads();
    }
  }
  // Run heads, ranked with a block scan over kSmall/kBlock items per thread.
  constexpr int kPer = kSmall / kBlock;
  uint32_t flags = 0, cnt = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    uint32_t p = tid * kPer + q;
    bool h = p < np && (p == 0 || s_row[p] != s_row[p - 1]);
    flags |= (h ? 1u : 0u) << q;
    cnt += h;
  }
  uint64_t total;
  uint64_t rank = block_exclusive_sum64_lds(cnt, s_scan, &total);
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    if ((flags >> q) & 1u) {
      uint32_t p = tid * kPer + q;
      uint64_t row = s_row[p];
      double sum = s_val[s_pos[p]];
      for (uint32_t r = p + 1; r < np && s_row[r] == row; ++r) sum += s_val[s_pos[r]];
      tmp_row[base + rank] = row;
      tmp_val[base + rank] = sum;
      ++rank;
    }
  }
  if (tid == 0) seg_u[s] = total;
  __syncthreads();  // LDS is reused by the next segment
  }
}

// 3a'. small segments, "hash" strategy: accumulate products of equal rows in
// an LDS open-addressing table (64-bit CAS on the row, FP64 LDS atomic add on
// the value), then bitonic-sort the table so that the unique rows come first.
// The summation order follows the atomics, so the last bits of a sum may vary
// from run to run (within the 1e-12 tolerance); indices do not.
constexpr int kHashSlots = 2 * kSmall;

__global__ __launch_bounds__(kBlock) void small_hash_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, const double *__restrict__ bv,
    const uint64_t *__restrict__ a_lo, const uint64_t *__restrict__ a_base,
    const uint64_t *__restrict__ prod_off, const uint64_t *__restrict__ seg_first, uint64_t nseg,
    uint64_t nb, const uint32_t *__restrict__ list2, const uint32_t *__restrict__ n2_dev, uint64_t *__restrict__ seg_u,
    uint64_t *__restrict__ tmp_row, double *__restrict__ tmp_val) {
  __shared__ unsigned long long s_key[kHashSlots];
  __shared__ double s_val[kHashSlots];
  __shared__ uint64_t s_scan[kBlock / kWave];
  constexpr unsigned long long kEmpty = ~0ull;
  const uint32_t tid = threadIdx.x;
  // Segments in a stride over the grid (their count is read on the device).
  const uint32_t n2 = *n2_dev;
  for (uint32_t it = blockIdx.x; it < n2; it += gridDim.x) {
  const uint64_t s = list2[it];
  const uint64_t first = seg_first[s];
  const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
  const uint64_t base = prod_off[first];
  const uint32_t np = (uint32_t)(prod_off[last] - base);
  uint32_t T = 2;
  while (T < 2 * np) T <<= 1;
  for (uint32_t i = tid; i < T; i += kBlock) {
    s_key[i] = kEmpty;
    s_val[i] = 0.0;
  }
  __syncthreads();
  for (uint32_t p = tid; p < np; p += kBlock) {
    uint64_t e, ai;
    locate(prod_off, first, last, p, a_lo, &e, &ai);
    unsigned long long row = ak[ai] - a_base[e];
    double val = bv[e] * av[ai];
    uint32_t slot = (uint32_t)((row * 0x9E3779B97F4A7C15ull) >> 40) & (T - 1);
    for (;;) {
      unsigned long long prev = atomicCAS(&s_key[slot], kEmpty, row);
      if (prev == kEmpty || prev == row) {
        atomicAdd(&s_val[slot], val);
        break;
      }
      slot = (slot + 1) & (T - 1);
    }
  }
  __syncthreads();
  for (uint32_t k = 2; k <= T; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = tid; i < T; i += kBlock) {
        uint32_t l = i ^ j;
        if (l > i) {
          unsigned long long ki = s_key[i], kl = s_key[l];
          bool up = (i & k) == 0;
          if ((ki > kl) == up && ki != kl) {
            double vi = s_val[i];
            s_key[i] = kl; s_key[l] = ki;
            s_val[i] = s_val[l]; s_val[l] = vi;
          }
        }
      }
      __syncthreads();
    }
  }
  uint32_t cnt = 0;
  for (uint32_t i = tid; i < T; i += kBlock) cnt += s_key[i] != kEmpty;
  uint64_t total;
  block_exclusive_sum64_lds(cnt, s_scan, &total);
  for (uint32_t i = tid; i < total; i += kBlock) {
    tmp_row[base + i] = s_key[i];
    tmp_val[base + i] = s_val[i];
  }
  if (tid == 0) seg_u[s] = total;
  __syncthreads();  // LDS is reused by the next segment
  }
}

// 3a''. dense accumulation (the sparse accumulator of csc_times, or of
// mtt_times over row ranks): one workgroup per segment with an accumulator of
// R <= kDenseRows doubles and a presence bitmap in LDS, indexed by the row (or
// by its rank among A's rows when RANKS).  The segment's B entries are taken
// in order; the rows of one A column are distinct, so the threads of one entry
// never meet on an index, and a barrier separates entries.  A row's first
// product is stored and later ones added: each sum is the same chain of
// additions, in ascending B row order, as the sort accumulator's, so the
// values are bit-identical to it.  Reading the bitmap in order gives the rows
// sorted, with no sort.
constexpr uint32_t kDenseRows = 8192;

template <bool RANKS>
__global__ __launch_bounds__(kBlock) void dense_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, const double *__restrict__ bv,
    const uint64_t *__restrict__ a_lo, const uint32_t *__restrict__ a_len, const uint64_t *__restrict__ a_base,
    const uint64_t *__restrict__ prod_off, const uint64_t *__restrict__ seg_first, uint64_t nseg, uint64_t nb,
    const uint32_t *__restrict__ list2, const uint32_t *__restrict__ n2_dev, uint32_t R, const uint32_t *__restrict__ arank,
    const uint32_t *__restrict__ rank_rows, uint64_t *__restrict__ seg_u, uint64_t *__restrict__ tmp_row,
    double *__restrict__ tmp_val) {
  __shared__ double s_acc[kDenseRows];
  __shared__ uint32_t s_bits[kDenseRows / 32];
  __shared__ uint32_t s_scan[kBlock / kWave];
  const uint32_t tid = threadIdx.x;
  // Segments in a stride over the grid (their count is read on the device).
  const uint32_t n2 = *n2_dev;
  for (uint32_t it = blockIdx.x; it < n2; it += gridDim.x) {
  const uint64_t s = list2[it];
  const uint64_t first = seg_first[s];
  const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
  const uint64_t base = prod_off[first];
  const uint32_t words = (R + 31) / 32;
  for (uint32_t w = tid; w < words; w += kBlock) s_bits[w] = 0;
  __syncthreads();
  for (uint64_t e = first; e < last; ++e) {
    const uint64_t lo = a_lo[e], ab = a_base[e];
    const uint32_t len = a_len[e];
    const double beta = bv[e];
    for (uint32_t i = tid; i < len; i += kBlock) {
      const uint32_t idx = RANKS ? arank[lo + i] : (uint32_t)(ak[lo + i] - ab);
      const double v = beta * av[lo + i];
      const uint32_t bit = 1u << (idx & 31);
      const uint32_t old = atomicOr(&s_bits[idx >> 5], bit);
      if (old & bit) s_acc[idx] += v;
      else s_acc[idx] = v;
    }
    __syncthreads();
  }
  // Thread t reads bitmap words t*per .. t*per + per - 1 (R <= 8192: per <= 1).
  constexpr uint32_t kPer = kDenseRows / 32 / kBlock;
  uint32_t mine[kPer], cnt = 0;
#pragma unroll
  for (uint32_t q = 0; q < kPer; ++q) {
    const uint32_t w = tid * kPer + q;
    mine[q] = w < words ? s_bits[w] : 0u;
    cnt += (uint32_t)__popc(mine[q]);
  }
  uint32_t total;
  uint32_t rank = block_exclusive_sum_bt<kBlock>(cnt, s_scan, &total);
#pragma unroll
  for (uint32_t q = 0; q < kPer; ++q) {
    uint32_t bits = mine[q];
    while (bits) {
      const uint32_t idx = (tid * kPer + q) * 32 + (uint32_t)(__ffs(bits) - 1);
      bits &= bits - 1;
      tmp_row[base + rank] = RANKS ? rank_rows[idx] : idx;
      tmp_val[base + rank] = s_acc[idx];
      ++rank;
    }
  }
  if (tid == 0) seg_u[s] = total;
  __syncthreads();  // LDS is reused by the next segment
  }
}

// 3b. large segments.  np3[r] = products of list3[r].
__global__ void large_sizes_kernel(const uint32_t *__restrict__ list3, uint32_t n3,
                                   const uint64_t *__restrict__ seg_first, uint64_t nseg, uint64_t nb,
                                   const uint64_t *__restrict__ prod_off, uint64_t *__restrict__ np3) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n3) return;
  uint64_t s = list3[r];
  uint64_t first = seg_first[s];
  uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
  np3[r] = prod_off[last] - prod_off[first];
}

// Outputs (or products) are handled in chunks of kChunk consecutive indices
// per workgroup round.  The offsets of the segments that cover a chunk are
// found with one global search per end and staged in LDS, where every index
// of the chunk finds its segment; a chunk covering more segments than fit
// (possible only with empty segments) searches the global offsets instead.
constexpr uint32_t kChunk = 2048;

struct ChunkSegs {
  uint64_t s0, s1;  // segments [s0, s1) cover the chunk
  bool staged;      // off[s0, s1) is in LDS
};

__device__ inline ChunkSegs stage_chunk(const uint64_t *__restrict__ off, uint64_t nseg, uint64_t q0, uint64_t q1,
                                        uint64_t *s_off, uint64_t *s_seg) {
  if (threadIdx.x == 0) {
    const uint64_t s0 = upper_index(off, 0, nseg, q0);
    s_seg[0] = s0;
    s_seg[1] = upper_index(off, s0, nseg, q1 - 1) + 1;
  }
  __syncthreads();
  ChunkSegs c{s_seg[0], s_seg[1], false};
  c.staged = c.s1 - c.s0 <= kChunk + 1;
  if (c.staged)
    for (uint64_t x = threadIdx.x; x < c.s1 - c.s0; x += blockDim.x) s_off[x] = off[c.s0 + x];
  __syncthreads();
  return c;
}

// The same, with the first covering segment known (s0, from chunk_starts_kernel)
// and the last bounded by s_hi: one short search instead of two long ones.
__device__ inline ChunkSegs stage_chunk_from(const uint64_t *__restrict__ off, uint64_t s0, uint64_t s_hi,
                                             uint64_t q1, uint64_t *s_off, uint64_t *s_seg) {
  if (threadIdx.x == 0) {
    s_seg[0] = s0;
    s_seg[1] = upper_index(off, s0, s_hi, q1 - 1) + 1;
  }
  __syncthreads();
  ChunkSegs c{s_seg[0], s_seg[1], false};
  c.staged = c.s1 - c.s0 <= kChunk + 1;
  if (c.staged)
    for (uint64_t x = threadIdx.x; x < c.s1 - c.s0; x += blockDim.x) s_off[x] = off[c.s0 + x];
  __syncthreads();
  return c;
}

// cs[k] = the segment holding index k * kChunk (k < nchunks), all chunks at
// once, so that no workgroup walks a long search chain alone.  nnz(C) is
// read on the device (off[nseg]), so the host need not wait for it.
// n_short: as in write_kernel (nothing to do when it points at zero).
__global__ void chunk_starts_kernel(const uint64_t *__restrict__ off, uint64_t nseg, uint64_t *__restrict__ cs,
                                    const uint32_t *__restrict__ n_short) {
  if (n_short && *n_short == 0) return;
  const uint64_t nchunks = (off[nseg] + kChunk - 1) / kChunk;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nchunks; k += (uint64_t)gridDim.x * blockDim.x)
    cs[k] = upper_index(off, 0, nseg, k * kChunk);
}

// Segment of index q and that segment's offset.
__device__ inline uint64_t chunk_segment(const ChunkSegs &c, const uint64_t *__restrict__ off, const uint64_t *s_off,
                                         uint64_t q, uint64_t *seg_off) {
  if (c.staged) {
    uint32_t lo = 0, hi = (uint32_t)(c.s1 - c.s0);
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_off[mid] <= q) lo = mid;
      else hi = mid;
    }
    *seg_off = s_off[lo];
    return c.s0 + lo;
  }
  const uint64_t s = upper_index(off, c.s0, c.s1, q);
  *seg_off = off[s];
  return s;
}

// Row ranks: the rows of A's entries marked in a bitmap over [0, h), the
// prefix of its words' popcounts, and the rows in rank order.  A row's rank is
// its position among the rows present, so ranks order products exactly as
// rows do while needing fewer key bits (the composite keys of the large
// segments can then take fewer radix passes, and 32-bit passes).
__global__ void row_bitmap_kernel(const uint64_t *__restrict__ ak, uint64_t na, Divider ah, uint32_t *__restrict__ bm) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < na; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = ak[e];
    const uint64_t row = key - divide(key, ah) * ah.d;
    atomicOr(&bm[row >> 5], 1u << (row & 31));
  }
}

__global__ void word_popc_kernel(const uint32_t *__restrict__ bm, uint64_t nw, uint32_t *__restrict__ cnt) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x)
    cnt[w] = (uint32_t)__popc(bm[w]);
}

__global__ void rank_rows_kernel(const uint32_t *__restrict__ bm, const uint64_t *__restrict__ pre, uint64_t nw,
                                 uint32_t *__restrict__ rows) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t bits = bm[w];
    uint64_t p = pre[w];
    while (bits) {
      rows[p++] = (uint32_t)(w * 32 + (uint32_t)(__ffs(bits) - 1));
      bits &= bits - 1;
    }
  }
}

// arank[e] = rank of the row of A entry e: computed once per A entry, then
// read by every product of that entry in place of its key.
__global__ void a_rank_kernel(const uint64_t *__restrict__ ak, uint64_t na, Divider ah,
                              const uint32_t *__restrict__ bm, const uint64_t *__restrict__ pre,
                              uint32_t *__restrict__ arank) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < na; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = ak[e];
    const uint64_t row = key - divide(key, ah) * ah.d;
    arank[e] = (uint32_t)(pre[row >> 5] + (uint64_t)__popc(bm[row >> 5] & ((1u << (row & 31)) - 1u)));
  }
}

// Products of the large segments with composite key r * width + row (or the
// row's rank when arank is set), kChunk per workgroup round.
__global__ __launch_bounds__(256) void large_expand_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, const double *__restrict__ bv,
    const uint64_t *__restrict__ a_lo, const uint64_t *__restrict__ a_base,
    const uint64_t *__restrict__ prod_off, const uint64_t *__restrict__ seg_first, uint64_t nseg,
    uint64_t nb, const uint32_t *__restrict__ list3, uint32_t n3, const uint64_t *__restrict__ ex_off,
    uint64_t total3, uint64_t width, const uint32_t *__restrict__ arank,
    uint64_t *__restrict__ key3, double *__restrict__ val3) {
  __shared__ uint64_t s_off[kChunk + 1];
  __shared__ uint64_t s_seg[2];
  for (uint64_t q0 = (uint64_t)blockIdx.x * kChunk; q0 < total3; q0 += (uint64_t)gridDim.x * kChunk) {
    const uint64_t q1 = min<uint64_t>(q0 + kChunk, total3);
    const ChunkSegs c = stage_chunk(ex_off, n3, q0, q1, s_off, s_seg);
    for (uint64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
      uint64_t base;
      const uint64_t r = chunk_segment(c, ex_off, s_off, q, &base);
      const uint64_t s = list3[r];
      const uint64_t first = seg_first[s];
      const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
      uint64_t e, ai;
      locate(prod_off, first, last, q - base, a_lo, &e, &ai);
      key3[q] = r * width + (arank ? (uint64_t)arank[ai] : ak[ai] - a_base[e]);
      val3[q] = bv[e] * av[ai];
    }
    __syncthreads();  // s_off is restaged next round
  }
}

__global__ void run_heads_kernel(const uint64_t *__restrict__ key3, uint64_t total3,
                                 uint32_t *__restrict__ heads) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; q < total3; q += (uint64_t)gridDim.x * blockDim.x)
    heads[q] = q == 0 || key3[q] != key3[q - 1];
}

// dst_base[r]: where the runs of large segment r start in tmp_row/tmp_val,
// less the global rank of its first run (mod 2^64), so that a run's slot is
// its global rank plus one per-segment word.
__global__ void large_base_kernel(const uint32_t *__restrict__ list3, uint32_t n3,
                                  const uint64_t *__restrict__ ex_off, const uint64_t *__restrict__ grank,
                                  const uint64_t *__restrict__ seg_first, const uint64_t *__restrict__ prod_off,
                                  uint64_t *__restrict__ dst_base) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n3) dst_base[r] = prod_off[seg_first[list3[r]]] - grank[ex_off[r]];
}

// Sums every run (in sorted = expansion order) and stores it in its segment's
// region of tmp_row/tmp_val.
__global__ void large_reduce_kernel(const uint64_t *__restrict__ key3, const double *__restrict__ val3,
                                    const uint32_t *__restrict__ heads, const uint64_t *__restrict__ grank,
                                    uint64_t total3, Divider width, const uint32_t *__restrict__ rows,
                                    const uint64_t *__restrict__ dst_base, uint64_t *__restrict__ tmp_row,
                                    double *__restrict__ tmp_val) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; q < total3; q += (uint64_t)gridDim.x * blockDim.x) {
    if (!heads[q]) continue;
    uint64_t key = key3[q];
    uint64_t r = divide(key, width);
    uint64_t row = key - r * width.d;
    if (rows) row = rows[row];
    double sum = val3[q];
    for (uint64_t x = q + 1; x < total3 && key3[x] == key; ++x) sum += val3[x];
    const uint64_t dst = grank[q] + dst_base[r];
    tmp_row[dst] = row;
    tmp_val[dst] = sum;
  }
}

// grank[total3] holds the number of runs (the scan's total).
__global__ void large_counts_kernel(const uint32_t *__restrict__ list3, uint32_t n3,
                                    const uint64_t *__restrict__ ex_off, const uint64_t *__restrict__ grank,
                                    uint64_t total3, uint64_t *__restrict__ seg_u) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n3) return;
  uint64_t lo = grank[ex_off[r]];
  uint64_t hi = grank[ex_off[r + 1] < total3 ? ex_off[r + 1] : total3];
  seg_u[list3[r]] = hi - lo;
}

// 3c. large segments cut into row pieces.  A segment's rows [rmin, rmax] are
// split into q equal ranges of about kPieceTarget products; every B entry's A
// column is sorted by row, so a piece is one contiguous run of each column,
// found by binary search.  One workgroup per piece loads its runs into LDS,
// merges them pairwise (ceil(log2 c) rounds, earlier B entries first on equal
// rows) into (row, B entry) order, and sums equal rows in that order: the
// same chain of additions, in ascending B row order, as the sort
// accumulator.  The pieces of a segment chain their
// unique-row counts (a look-back on workgroup tickets, so a piece waits only
// for pieces already running), and each writes its rows right after its
// predecessor's.  A piece of more than kPieceCap products (rows clustered in
// one range) goes through in row sub-ranges found by halving; a segment of
// more than kPieceLists B entries, or rows spanning more than 2^32 per piece,
// takes the global sort instead.
constexpr uint32_t kPieceBlock = 512;
constexpr uint32_t kPieceCap = 3584;  // 7 per thread; 48 KB of LDS: three workgroups (24 waves) per CU
constexpr uint32_t kPieceTarget = 3072;
constexpr uint32_t kPieceLists = 64;
constexpr uint32_t kPieceItems = kPieceCap / kPieceBlock;
// Status word of a piece: [63:62] 1 = its own unique count, 2 = the count up
// to and including it within its segment; [61:0] the count.
constexpr uint64_t kPieceAgg = uint64_t(1) << 62;
constexpr uint64_t kPieceIncl = uint64_t(2) << 62;
constexpr uint64_t kPieceValue = kPieceAgg - 1;

// Per large segment r: rmin, piece width, and (q << 32 | q * c) for the scan
// of piece and (piece, B entry) pair offsets.  q = 0 marks a segment for the
// global sort.  One wave per segment, lane i on B entry i; the number of
// large segments is read on the device (the kernel is launched before the
// host knows it).
__global__ __launch_bounds__(256) void piece_plan_kernel(
    const uint64_t *__restrict__ ak, const uint64_t *__restrict__ a_lo, const uint32_t *__restrict__ a_len,
    const uint64_t *__restrict__ a_base, const uint64_t *__restrict__ prod_off,
    const uint64_t *__restrict__ seg_first, uint64_t nseg, uint64_t nb, const uint32_t *__restrict__ list3,
    const uint64_t *__restrict__ n3_dev, uint64_t n3_ub, uint64_t *__restrict__ rmin, uint64_t *__restrict__ width,
    uint64_t *__restrict__ qc, uint32_t *__restrict__ ovf, uint32_t *__restrict__ ovf_count,
    uint64_t *__restrict__ seg_pos) {
  const uint64_t n3 = *n3_dev;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  // qc is zero from n3 up to the host's bound n3_ub: the scans and searches
  // over qc may run to the bound when the host has not read n3.
  for (uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave + n3; r < n3_ub; r += waves)
    if (lane == 0) qc[r] = 0;
  for (uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < n3; r += waves) {
    const uint64_t s = list3[r];
    const uint64_t first = seg_first[s];
    const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
    const uint64_t c = last - first, np = prod_off[last] - prod_off[first];
    uint64_t q = 0, lo = ~0ull, hi = 0, w = 0;
    if (c <= kPieceLists) {
      if (lane < c) {
        const uint64_t e = first + lane;
        const uint32_t n = a_len[e];
        if (n) {
          lo = ak[a_lo[e]] - a_base[e];
          hi = ak[a_lo[e] + n - 1] - a_base[e];
        }
      }
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint64_t ol = __shfl_xor(lo, off, kWave), oh = __shfl_xor(hi, off, kWave);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
      }
      // np > 0, so some column is non-empty and lo <= hi.
      const uint64_t range = hi - lo + 1;
      q = (np + kPieceTarget - 1) / kPieceTarget;
      w = (range + q - 1) / q;
      if (w > 0xFFFFFFFFull) q = 0;
      else q = (range + w - 1) / w;
    }
    if (lane == 0) {
      seg_pos[2 * r] = first;
      seg_pos[2 * r + 1] = prod_off[first];
      rmin[r] = lo;
      width[r] = w;
      ovf[r] = q == 0;
      if (q == 0) atomicAdd(ovf_count, 1u);
      qc[r] = (q << 32) | (q * c);
    }
  }
}

// Last r in [0, n3) with part(off[r]) <= t (part: the high or low 32 bits).
template <bool HIGH>
__device__ inline uint32_t piece_segment(const uint64_t *__restrict__ off, uint32_t n3, uint64_t t) {
  uint32_t lo = 0, hi = n3;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint64_t v = HIGH ? off[mid] >> 32 : off[mid] & 0xFFFFFFFFull;
    if (v <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// bounds[pair]: first A entry of piece p in the column of B entry i (pair =
// the segment's pair offset + p * c + i).
__global__ void piece_bounds_kernel(const uint64_t *__restrict__ ak, const uint64_t *__restrict__ a_lo,
                                    const uint32_t *__restrict__ a_len, const uint64_t *__restrict__ a_base,
                                    const uint64_t *__restrict__ seg_first, uint64_t nseg, uint64_t nb,
                                    const uint32_t *__restrict__ list3, uint32_t n3,
                                    const uint64_t *__restrict__ off, const uint64_t *__restrict__ rmin,
                                    const uint64_t *__restrict__ width, uint64_t *__restrict__ bounds,
                                    uint32_t *__restrict__ piece_seg) {
  const uint64_t total = off[n3] & 0xFFFFFFFFull;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const uint32_t r = piece_segment<false>(off, n3, t);
    const uint64_t s = list3[r];
    const uint64_t first = seg_first[s];
    const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
    const uint64_t c = last - first, local = t - (off[r] & 0xFFFFFFFFull);
    const uint64_t p = local / c, e = first + (local - p * c);
    if (e == first) piece_seg[(off[r] >> 32) + p] = r;
    const uint64_t key = a_base[e] + rmin[r] + p * width[r];
    uint64_t lo = a_lo[e], hi = lo + a_len[e];
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (ak[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    bounds[t] = lo;
  }
}

__global__ __launch_bounds__(kPieceBlock, 6) void piece_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, const double *__restrict__ bv,
    const uint64_t *__restrict__ a_lo, const uint32_t *__restrict__ a_len, const uint64_t *__restrict__ a_base,
    const uint32_t *__restrict__ list3, uint32_t n3, const uint64_t *__restrict__ off,
    const uint64_t *__restrict__ rmin, const uint64_t *__restrict__ width, const uint64_t *__restrict__ bounds,
    const uint32_t *__restrict__ piece_seg, const uint64_t *__restrict__ seg_pos, uint32_t *__restrict__ ticket,
    uint64_t *__restrict__ status, uint64_t *__restrict__ seg_u, uint64_t *__restrict__ tmp_row,
    double *__restrict__ tmp_val, int direct, const uint64_t *__restrict__ out_part,
    const uint64_t *__restrict__ colid, uint64_t ah, uint64_t *__restrict__ ck, double *__restrict__ cv) {
  // Rows (offsets from the piece's first row) and load positions, in two
  // buffers for the merge rounds, and the B entry of each load position.  The
  // products are formed again from A (in cache) when the rows are summed,
  // which keeps the LDS at three workgroups per CU.
  __shared__ uint32_t s_row[2][kPieceCap];
  __shared__ uint16_t s_idx[2][kPieceCap];
  __shared__ uint8_t s_run[kPieceCap];
  __shared__ uint32_t s_off[kPieceLists + 1];
  __shared__ uint64_t s_src[kPieceLists];   // first A entry of each run of the current range
  __shared__ uint64_t s_next[kPieceLists];  // first A entry past the current range
  __shared__ uint64_t s_end[kPieceLists];   // first A entry past the piece
  __shared__ uint64_t s_key0[kPieceLists];
  __shared__ double s_b[kPieceLists];
  __shared__ uint64_t s_scan[kPieceBlock / kWave];
  __shared__ uint32_t s_t;
  __shared__ uint64_t s_prefix, s_lo;
  const uint32_t tid = threadIdx.x;
  // The grid may be sized by a bound: workgroups past the number of pieces
  // leave before taking a ticket (every ticket below it is taken).
  const uint64_t qtot = off[n3] >> 32;
  if (blockIdx.x >= qtot) return;
  if (tid == 0) s_t = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint64_t t = s_t;
  // The piece's segment and its places: loads that do not wait on one
  // another, so the start-up latency is a few round trips.
  const uint32_t r = piece_seg[t];  // piece_seg holds q_ub >= grid entries
  if (t >= qtot) return;
  const uint64_t s = list3[r], first = seg_pos[2 * r], seg_base = seg_pos[2 * r + 1];
  const uint64_t o0 = off[r], o1 = off[r + 1];
  const uint32_t c = (uint32_t)(((o1 & 0xFFFFFFFFull) - (o0 & 0xFFFFFFFFull)) / ((o1 >> 32) - (o0 >> 32)));
  const uint64_t q = (o1 >> 32) - (o0 >> 32), p = t - (o0 >> 32);
  const uint64_t pb = (o0 & 0xFFFFFFFFull) + p * c;
  const uint64_t w = width[r], v0 = rmin[r] + p * w;
  // Direct mode: rows go straight into C, at the segment's offset among the
  // other segments plus the count of every earlier piece (the look-back then
  // runs over all pieces, in output order), keyed by the output column.
  const uint64_t dst_base = direct ? out_part[s] : seg_base;
  const uint64_t key_add = direct ? ah * colid[first] : 0;
  uint64_t *const row_out = direct ? ck : tmp_row;
  double *const val_out = direct ? cv : tmp_val;
  uint64_t total_all;
  if (tid < kWave) {  // c <= kPieceLists = one wave: lane i takes B entry i
    uint64_t len = 0;
    if (tid < c) {
      const uint64_t e = first + tid;
      const uint64_t b0 = bounds[pb + tid];
      const uint64_t b1 = p + 1 < q ? bounds[pb + c + tid] : a_lo[e] + a_len[e];
      len = b1 - b0;
      s_src[tid] = b0;
      s_next[tid] = b1;
      s_end[tid] = b1;
      s_key0[tid] = a_base[e] + v0;
      s_b[tid] = bv[e];
    }
    const uint64_t incl = wave_inclusive_sum64(len);
    if (tid < c) s_off[tid] = (uint32_t)min<uint64_t>(incl - len, kPieceCap + 1);
    if (tid == kWave - 1) {
      s_off[c] = (uint32_t)min<uint64_t>(incl, kPieceCap + 1);
      s_scan[0] = incl;
    }
  }
  __syncthreads();
  total_all = s_scan[0];
  __syncthreads();
  // Run (B entry) of position x: the last run starting at or before x.
  auto run_of = [&](uint32_t x) {
    uint32_t lo = 0, hi = c;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_off[mid] <= x) lo = mid;
      else hi = mid;
    }
    return lo;
  };
  uint32_t cur = 0;
  // Loads the runs s_src[i] + [0, s_off[i+1] - s_off[i]) (at most kPieceCap
  // products), orders them by (row, B entry), marks run heads (s_head) and
  // returns the number of distinct rows; `cur` names the ordered buffer.
  auto order_rows = [&]() -> uint64_t {
    const uint32_t total = s_off[c];
    for (uint32_t x = tid; x < total; x += kPieceBlock) {
      const uint32_t i = run_of(x);
      const uint64_t ai = s_src[i] + (x - s_off[i]);
      s_row[0][x] = (uint32_t)(ak[ai] - s_key0[i]);
      s_idx[0][x] = (uint16_t)x;
      s_run[x] = (uint8_t)i;
    }
    __syncthreads();
    // Pairwise merge rounds: after the round of width mw, every run holds
    // 2mw consecutive B entries.  Each thread writes one contiguous stretch
    // of output positions, locating it in its pair by a merge-path search; on
    // equal rows the left run (earlier B entries) goes first.
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
        // The two run heads stay in registers: one LDS read per output.
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
    // Run heads over `per` consecutive positions per thread (per <= kPieceItems).
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
    uint32_t *s_head = s_row[cur ^ 1];  // free after the last merge round
    for (uint32_t x = h0; x < h1; ++x)
      if ((flags >> (x - h0)) & 1u) s_head[urank++] = x;
    __syncthreads();
    return utotal;
  };
  // The ordered rows' sums, kept in registers (row u = tid + k * kPieceBlock
  // in slot k) so that they are gathered before the look-back waits; the rows
  // themselves stay in LDS (head list, ordered rows) until they are stored.
  double hsum[kPieceItems];
  auto sum_rows = [&](uint64_t utotal) {
    // Each product rounded on its own, never fused with the addition that
    // follows (as in every other accumulator).  The first product of every
    // row is fetched for all of a thread's rows before any is summed, so the
    // fetches overlap.
#pragma clang fp contract(off)
    const uint32_t total = s_off[c];
    const uint32_t *krow = s_row[cur], *s_head = s_row[cur ^ 1];
    const uint16_t *kidx = s_idx[cur];
    auto product = [&](uint32_t y) {
      const uint32_t x = kidx[y], i = s_run[x];
      return s_b[i] * av[s_src[i] + (x - s_off[i])];
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
  // Stores the summed rows at dst0 (their positions in row_out/val_out).
  auto store_rows = [&](uint64_t dst0, uint64_t utotal) {
    const uint32_t nu = (uint32_t)((utotal + kPieceBlock - 1) / kPieceBlock);
    const uint32_t *krow = s_row[cur], *s_head = s_row[cur ^ 1];
#pragma unroll
    for (uint32_t k = 0; k < kPieceItems; ++k) {
      const uint32_t u = tid + k * kPieceBlock;
      if (k < nu && u < utotal) {
        row_out[dst0 + u] = v0 + krow[s_head[u]] + key_add;
        val_out[dst0 + u] = hsum[k];
      }
    }
  };
  // Decoupled look-back within the segment (over all pieces in direct mode):
  // a piece publishes its own count as soon as it has it (an oversized piece
  // cannot; the first piece publishes it as inclusive), gathers its sums,
  // then wave 0 sums its predecessors' words, 64 per round trip, back to
  // the first inclusive one.  Only counts travel between pieces (the rows are read by a
  // later kernel), so relaxed atomics suffice: no cache write-back or
  // invalidate.  Sets s_prefix.
  const bool first_piece = direct ? t == 0 : p == 0;
  auto look_back = [&]() {
    if (tid < kWave) {
      uint64_t prefix = 0;
      if (!first_piece) {
        const int64_t seg0 = direct ? 0 : (int64_t)(t - p);
        int64_t base = (int64_t)t - 1;
        for (;;) {
          const int64_t idx = base - (int64_t)tid;
          uint64_t wd = 0;
          bool ready = true, term = true;
          if (idx >= seg0) {
            wd = status_load(&status[idx]);
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
  uint64_t count;
  if (total_all <= kPieceCap) {
    count = order_rows();
    if (tid == 0) status_store(&status[t], (first_piece ? kPieceIncl : kPieceAgg) | count);
    sum_rows(count);
    look_back();
    if (tid == 0 && !first_piece) status_store(&status[t], kPieceIncl | (s_prefix + count));
    store_rows(dst_base + s_prefix, count);
  } else {
    // Rows clustered: the piece goes through in row sub-ranges of at most
    // kPieceCap products, each found by halving the remaining range (a single
    // row holds at most c <= kPieceLists products, so the halving ends), one
    // after another, each written after the last.  Its successors wait for
    // its inclusive count.
    look_back();
    count = 0;
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
              if (ak[mid] < key) lo = mid + 1;
              else hi = mid;
            }
            nxt = lo;
            len = lo - s_src[tid];
          }
          const uint64_t tot = __shfl(wave_inclusive_sum64(len), kWave - 1, kWave);
          if (tot <= kPieceCap || span == 1) break;
          span = (span + 1) / 2;
        }
        uint64_t len = tid < c ? nxt - s_src[tid] : 0;
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
      sum_rows(u);
      store_rows(dst_base + s_prefix + count, u);
      count += u;
      __syncthreads();
      if (tid < c) s_src[tid] = s_next[tid];
      __syncthreads();
    }
    if (tid == 0) status_store(&status[t], kPieceIncl | (s_prefix + count));
  }
  if (!direct && tid == 0 && p + 1 == q) seg_u[s] = s_prefix + count;
}

// Direct mode: each pieced segment's row count (the difference of the
// inclusive counts that close it and its predecessor), and its mark for the
// output kernel, which skips it.
// n3_host: the number of long segments or the host's bound on it; n3_dev:
// the number on the device, read when the host has only the bound (the
// global sort of fallback segments reuses that word, but it runs only when
// the host has read the number).
__global__ void piece_total_kernel(const uint32_t *__restrict__ list3, uint32_t n3_host,
                                   const uint64_t *__restrict__ n3_dev, const uint64_t *__restrict__ off,
                                   const uint64_t *__restrict__ status, uint64_t *__restrict__ seg_u,
                                   uint8_t *__restrict__ direct) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n3_host || (n3_dev && r >= *n3_dev)) return;
  const uint64_t t0 = off[r] >> 32, t1 = off[r + 1] >> 32;
  if (t1 == t0) return;  // taken by the global sort
  const uint64_t hi = status[t1 - 1] & kPieceValue, lo = t0 ? status[t0 - 1] & kPieceValue : 0;
  seg_u[list3[r]] = hi - lo;
  direct[list3[r]] = 1;
}

// The segments whose pieces gave up, for the global sort.
__global__ void piece_fallback_kernel(const uint32_t *__restrict__ list3, uint32_t n3,
                                      const uint32_t *__restrict__ ovf, uint32_t *__restrict__ out,
                                      uint32_t *__restrict__ count) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n3 && ovf[r]) out[atomicAdd(count, 1u)] = list3[r];
}

__global__ void fill_words_kernel(uint64_t *__restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// 4a. output of the single-entry segments (C(:,j) = b * A(:,k)), when they
// are long on average.  Every A column is read by many such segments: the
// segments are dealt to the XCDs by where their A run starts (the x-th eighth
// of A to XCD x; workgroups are dealt to the XCDs in turn), so that each
// XCD's L2 holds the eighth of A it reads instead of every XCD streaming all
// of A.  A workgroup takes 64 consecutive segments (lane j: segment g*64 + j
// when it is single and in the XCD's eighth) and its waves take turns over
// chunks of 64*kCopyItems of their products; each product finds its segment
// by walking the segments that meet its chunk (a ballot, then one compare
// per item and segment) and takes that segment's place by lane shuffles.
// Every load of a chunk is issued before its stores.  Correctness does not
// depend on the placement.
constexpr uint32_t kCopyItems = 8;

__device__ inline uint64_t readlane64(uint64_t v, uint32_t j) {
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, (int)j);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)j);
  return ((uint64_t)hi << 32) | lo;
}

// The grid is a multiple of 8.
__global__ __launch_bounds__(256) void single_copy_kernel(
    const uint64_t *__restrict__ ak, const double *__restrict__ av, uint64_t na, const double *__restrict__ bv,
    const uint64_t *__restrict__ a_lo, const uint32_t *__restrict__ a_len, const uint64_t *__restrict__ a_base,
    const uint64_t *__restrict__ colid, const uint64_t *__restrict__ seg_first, uint64_t nseg, uint64_t nb,
    const uint64_t *__restrict__ out_off, uint64_t ah, uint64_t *__restrict__ ck, double *__restrict__ cv) {
  const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint32_t x = blockIdx.x & 7;
  const uint64_t ngroups = (nseg + kWave - 1) / kWave;
  for (uint64_t g = blockIdx.x >> 3; g < ngroups; g += gridDim.x >> 3) {
    const uint64_t s = g * kWave + lane;
    uint64_t n = 0, o = 0, lo = 0, kshift = 0;
    double b = 0.0;
    if (s < nseg) {
      const uint64_t first = seg_first[s];
      const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
      if (last - first == 1) {
        lo = a_lo[first];
        if (lo * 8 / na == x) {
          n = a_len[first];
          o = out_off[s];
          kshift = ah * colid[first] - a_base[first];
          b = bv[first];
        }
      }
    }
    const uint64_t incl = wave_inclusive_sum64(n), excl = incl - n;
    const uint64_t ashift = lo - excl, oshift = o - excl;  // A index and output index - product number
    const uint64_t total = __shfl(incl, kWave - 1, kWave);
    constexpr uint64_t kSpan = (uint64_t)kWave * kCopyItems;
    for (uint64_t p0 = (uint64_t)w * kSpan; p0 < total; p0 += 4 * kSpan) {
      const uint64_t pend = min<uint64_t>(p0 + kSpan, total);
      // The segments that meet [p0, pend), ascending: the first holds p0, so
      // every item gets a segment; later ones take the items past their start.
      uint64_t m = __ballot(n > 0 && excl < pend && incl > p0);
      uint32_t js[kCopyItems];
#pragma unroll
      for (uint32_t u = 0; u < kCopyItems; ++u) js[u] = 0;
      while (m) {
        const uint32_t j = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const uint64_t ej = readlane64(excl, j);
#pragma unroll
        for (uint32_t u = 0; u < kCopyItems; ++u)
          if (p0 + u * kWave + lane >= ej) js[u] = j;
      }
      uint64_t key[kCopyItems], dst[kCopyItems];
      double val[kCopyItems];
#pragma unroll
      for (uint32_t u = 0; u < kCopyItems; ++u) {
        const uint64_t p = p0 + u * kWave + lane;
        const uint64_t pc = p < total ? p : total - 1;  // a place inside the last segment
        // Every lane shuffles: a lane reading from an inactive lane gets nothing.
        const uint64_t ai = __shfl(ashift, (int)js[u], kWave) + pc;
        const uint64_t os = __shfl(oshift, (int)js[u], kWave);
        key[u] = ak[ai] + __shfl(kshift, (int)js[u], kWave);
        val[u] = __shfl(b, (int)js[u], kWave) * av[ai];
        dst[u] = p < total ? os + p : ~0ull;
      }
#pragma unroll
      for (uint32_t u = 0; u < kCopyItems; ++u) {
        if (dst[u] != ~0ull) {
          ck[dst[u]] = key[u];
          cv[dst[u]] = val[u];
        }
      }
    }
  }
}

// 4. output: the elements of C in order, kChunk per workgroup round.
__global__ __launch_bounds__(256) void write_kernel(const uint64_t *__restrict__ ak, const double *__restrict__ av,
                             const double *__restrict__ bv, const uint64_t *__restrict__ a_lo,
                             const uint64_t *__restrict__ a_base, const uint64_t *__restrict__ colid,
                             const uint64_t *__restrict__ prod_off, const uint64_t *__restrict__ seg_first,
                             uint64_t nseg, uint64_t nb, const uint64_t *__restrict__ out_off,
                             const uint64_t *__restrict__ tmp_row, const double *__restrict__ tmp_val,
                             uint64_t ah, const uint64_t *__restrict__ cs,
                             const uint8_t *__restrict__ direct, const uint32_t *__restrict__ n_short,
                             uint64_t *__restrict__ ck, double *__restrict__ cv) {
  __shared__ uint64_t s_off[kChunk + 1];
  __shared__ uint64_t s_seg[2];
  // Only short segments left to write (every other one is marked), and none
  // of them: nothing to do.
  if (n_short && *n_short == 0) return;
  const uint64_t nc = out_off[nseg];  // nnz(C), read on the device
  const uint64_t nchunks = (nc + kChunk - 1) / kChunk;
  for (uint64_t k = blockIdx.x; k < nchunks; k += gridDim.x) {
    const uint64_t o0 = k * kChunk, o1 = min<uint64_t>(o0 + kChunk, nc);
    const ChunkSegs c = stage_chunk_from(out_off, cs[k], k + 1 < nchunks ? min<uint64_t>(cs[k + 1] + 1, nseg) : nseg,
                                         o1, s_off, s_seg);
    // kPer elements per thread, all loads issued before any store, so that
    // their chains of dependent loads overlap.
    constexpr uint32_t kPer = kChunk / 256;
    // Segments whose rows the pieces wrote are skipped.
    uint64_t key[kPer];
    double val[kPer];
    uint32_t skip = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint64_t o = o0 + threadIdx.x + k * 256;
      if (o < o1) {
        uint64_t base;
        const uint64_t s = chunk_segment(c, out_off, s_off, o, &base);
        if (direct[s]) {
          skip |= 1u << k;
          continue;
        }
        const uint64_t q = o - base;
        const uint64_t first = seg_first[s];
        const uint64_t last = s + 1 < nseg ? seg_first[s + 1] : nb;
        if (last - first == 1) {
          const uint64_t ai = a_lo[first] + q;
          key[k] = ak[ai] - a_base[first] + ah * colid[first];
          val[k] = bv[first] * av[ai];
        } else {
          key[k] = tmp_row[prod_off[first] + q] + ah * colid[first];
          val[k] = tmp_val[prod_off[first] + q];
        }
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint64_t o = o0 + threadIdx.x + k * 256;
      if (o < o1 && !((skip >> k) & 1u)) {
        ck[o] = key[k];
        cv[o] = val[k];
      }
    }
    __syncthreads();  // s_off is restaged next round
  }
}

inline uint32_t chunk_grid(uint64_t n) {
  const uint64_t b = (n + kChunk - 1) / kChunk;
  return (uint32_t)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

inline uint32_t grid_for(uint64_t n, uint32_t cap = 65536) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  return (uint32_t)(b > cap ? cap : b);
}

struct ValidationOff {
  lnt_context ctx;
  int saved;
  explicit ValidationOff(lnt_context c) : ctx(c), saved(c->validate) { ctx->validate = 0; }
  ~ValidationOff() { ctx->validate = saved; }
};

struct Carve {
  char *p;
  template <typename T>
  T *take(uint64_t count) {
    T *r = (T *)p;
    p += ((count * sizeof(T) + 255) / 256) * 256;
    return r;
  }
};

// Two device words, one stream synchronisation.
int read_two_u64(lnt_context ctx, const uint64_t *a, const uint64_t *b, uint64_t out[2]) {
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 8, a, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 10, b, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(out, ctx->pinned + 8, 2 * sizeof(uint64_t));
  return LNT_OK;
}

int read_u64(lnt_context ctx, const uint64_t *dev, uint64_t *out, int count) {
  LNT_HIP(hipMemcpyAsync(ctx->pinned + 8, dev, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, ctx->stream));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(out, ctx->pinned + 8, sizeof(uint64_t) * count);
  return LNT_OK;
}

}  // namespace

static int check_lattice(const int64_t a_dims[3], const int64_t b_dims[3]) {
  for (int i = 0; i < 3; ++i)
    if (a_dims[i] < 1 || b_dims[i] < 1) return fail(LNT_ERR_INVALID_ARGUMENT, "lattice dimensions must be >= 1");
  if (a_dims[1] != b_dims[0] || a_dims[2] != b_dims[2])
    return fail(LNT_ERR_DIMENSION_MISMATCH, "A's width and depth must equal B's height and depth");
  unsigned __int128 a = (unsigned __int128)a_dims[0] * a_dims[1] * a_dims[2];
  unsigned __int128 b = (unsigned __int128)b_dims[0] * b_dims[1] * b_dims[2];
  unsigned __int128 c = (unsigned __int128)a_dims[0] * b_dims[1] * b_dims[2];
  const unsigned __int128 lim = (unsigned __int128)1 << 63;
  if (a > lim || b > lim || c > lim) return fail(LNT_ERR_TOO_LARGE, "lattice dimensionality exceeds 2^63");
  return LNT_OK;
}

// Row-piece arrays: the plan per large segment (sized by the segment count),
// then per piece and per (piece, B entry) pair.
struct PieceArrays {
  uint64_t *rmin, *width, *qc, *qc_off, *seg_pos, *bounds, *status;
  uint32_t *ovf, *list, *piece_seg;
};

struct EntryArrays {
  uint64_t *a_lo, *a_base, *colid, *prod_off, *seg_idx, *seg_first, *seg_u, *out_off, *np3, *ex_off;
  uint32_t *a_len, *head, *list2, *list3, *counts;
  uint8_t *direct;  // segments whose rows the pieces write straight into C
  size_t scan_words;
};

// Largest A column count (w*d) for which the CSC column-pointer array is
// built; beyond it the lookup searches A's keys (the DCSC form).
constexpr uint64_t kMaxPtrColumns = uint64_t(1) << 26;

static int entry_phase(lnt_context ctx, const uint64_t *ak, uint64_t na, const int64_t a_dims[3],
                       const uint64_t *bk, uint64_t nb, const int64_t b_dims[3], bool csc, EntryArrays *E) {
  size_t scan_words = scan_workspace_words(nb + 1);
  size_t bytes = (nb + 1) * (8 * 10 + 4 * 4 + 1) + 17 * 256 + 64 + scan_words * 8 + 8192;
  LNT_TRY(ensure(ctx->spg_ws, bytes));
  Carve c{(char *)ctx->spg_ws.ptr};
  E->a_lo = c.take<uint64_t>(nb + 1);
  E->a_base = c.take<uint64_t>(nb + 1);
  E->colid = c.take<uint64_t>(nb + 1);
  E->prod_off = c.take<uint64_t>(nb + 1);
  E->seg_idx = c.take<uint64_t>(nb + 1);
  E->seg_first = c.take<uint64_t>(nb + 1);
  E->seg_u = c.take<uint64_t>(nb + 1);
  E->out_off = c.take<uint64_t>(nb + 1);
  E->np3 = c.take<uint64_t>(nb + 1);
  E->ex_off = c.take<uint64_t>(nb + 1);
  E->a_len = c.take<uint32_t>(nb + 1);
  E->head = c.take<uint32_t>(nb + 1);
  E->list2 = c.take<uint32_t>(nb + 1);
  E->list3 = c.take<uint32_t>(nb + 1);
  E->counts = c.take<uint32_t>(16);
  E->direct = c.take<uint8_t>(nb + 1);
  E->scan_words = scan_words;
  LatticeDims L;
  L.ah = (uint64_t)a_dims[0];
  L.aw = (uint64_t)a_dims[1];
  L.bw = (uint64_t)b_dims[1];
  L.b_tab = make_divider((uint64_t)b_dims[0] * (uint64_t)b_dims[1]);
  L.b_h = make_divider((uint64_t)b_dims[0]);
  const uint64_t *cp = nullptr;
  const uint64_t ncols = (uint64_t)a_dims[1] * (uint64_t)a_dims[2];
  if (csc && ncols <= kMaxPtrColumns) {
    LNT_TRY(ensure(ctx->spg_cp, (ncols + 1) * 8));
    col_ptr_kernel<<<grid_for(ncols + 1), 256, 0, ctx->stream>>>(ak, na, L.ah, ncols, (uint64_t *)ctx->spg_cp.ptr);
    LNT_HIP(hipGetLastError());
    cp = (const uint64_t *)ctx->spg_cp.ptr;
  }
  entry_kernel<<<grid_for(nb), 256, 0, ctx->stream>>>(ak, na, bk, nb, L, cp, E->a_lo, E->a_len, E->a_base,
                                                        E->colid, E->head);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

// Ranks of A's rows among the rows present (the row compression of
// mtt_times, SparseLatticeBase.h:1697-1716, over the whole lattice: the map is
// monotone, so per-tab and global ranks order every column's rows alike).
struct RankTables {
  uint32_t *bm = nullptr, *cnt = nullptr, *rows = nullptr, *ar = nullptr;
  uint64_t *pre = nullptr;
  uint64_t nw = 0;
};

// Bitmap of the rows present and the prefix of its word popcounts;
// *nrank = the number of distinct rows.  Needs h <= 2^31.
static int rank_count(lnt_context ctx, const uint64_t *ak, uint64_t na, uint64_t ah, size_t scan_words,
                      RankTables *t, uint64_t *nrank) {
  const uint64_t nw = (ah + 31) / 32;
  LNT_TRY(ensure(ctx->spg_rank, nw * 8 + (nw + 1) * 8 + na * 8 + 1024));
  Carve c{(char *)ctx->spg_rank.ptr};
  t->nw = nw;
  t->bm = c.take<uint32_t>(nw);
  t->cnt = c.take<uint32_t>(nw);
  t->pre = c.take<uint64_t>(nw + 1);
  t->rows = c.take<uint32_t>(na);
  t->ar = c.take<uint32_t>(na);
  LNT_TRY(ensure(ctx->aux, std::max(scan_words, scan_workspace_words(nw)) * 8 + 64));
  LNT_TRY(launch_fill_u32(ctx, t->bm, 0, nw));
  row_bitmap_kernel<<<grid_for(na), 256, 0, ctx->stream>>>(ak, na, make_divider(ah), t->bm);
  word_popc_kernel<<<grid_for(nw), 256, 0, ctx->stream>>>(t->bm, nw, t->cnt);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, t->cnt, t->pre, nw, t->pre + nw, ctx->aux, 0));
  return read_u64(ctx, t->pre + nw, nrank, 1);
}

// rows[rank] = row, ar[e] = rank of A entry e's row (after rank_count).
static int rank_tables(lnt_context ctx, const uint64_t *ak, uint64_t na, uint64_t ah, const RankTables &t) {
  rank_rows_kernel<<<grid_for(t.nw), 256, 0, ctx->stream>>>(t.bm, t.pre, t.nw, t.rows);
  a_rank_kernel<<<grid_for(na), 256, 0, ctx->stream>>>(ak, na, make_divider(ah), t.bm, t.pre, t.ar);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

int lattice_nprod_device(lnt_context ctx, const uint64_t *ak, uint64_t na, const int64_t a_dims[3],
                         const uint64_t *bk, uint64_t nb, const int64_t b_dims[3], uint64_t *nprod_host) {
  LNT_TRY(check_lattice(a_dims, b_dims));
  *nprod_host = 0;
  if (na == 0 || nb == 0) return LNT_OK;
  EntryArrays E;
  LNT_TRY(entry_phase(ctx, ak, na, a_dims, bk, nb, b_dims, false, &E));
  // The scan workspace lives after out_off's region in the same carve: reuse
  // ctx->aux for it.
  LNT_TRY(ensure(ctx->aux, E.scan_words * 8 + 64));
  LNT_TRY(exclusive_scan<uint32_t>(ctx, E.a_len, E.prod_off, nb, E.prod_off + nb, ctx->aux, 0));
  return read_u64(ctx, E.prod_off + nb, nprod_host, 1);
}

namespace {

// lattice_mult_impl's answer when it ran without waiting for the segment
// lists and found segments the row pieces cannot take: run again, waiting.
constexpr int kRunAgainWaiting = 1 << 20;

// wait_for_lists: read the number of short, long and fallback segments on
// the host before launching their kernels (otherwise the launches are sized
// by bounds and read the numbers on the device, and a fallback segment --
// more than kPieceLists B entries, or rows too far apart -- found at the end
// asks for a second run that waits).
int lattice_mult_impl(lnt_context ctx, const uint64_t *ak, const double *av, uint64_t na, const int64_t a_dims[3],
                      const uint64_t *bk, const double *bv, uint64_t nb, const int64_t b_dims[3], int strategy,
                      uint64_t *ck, double *cv, uint64_t capacity, uint64_t *nc_host, uint64_t *nprod_host,
                      uint64_t **ck_alloc, double **cv_alloc, bool wait_for_lists) {
  LNT_TRY(check_lattice(a_dims, b_dims));
  *nc_host = 0;
  if (nprod_host) *nprod_host = 0;
  if (na == 0 || nb == 0) return LNT_OK;
  const uint64_t ah = (uint64_t)a_dims[0];
  const int acc = strategy & ~LNT_SPGEMM_CSC;
  int slot = prof_begin(ctx, kProfLatticeMult, 16.0 * ((double)na + nb));
  EntryArrays E;
  LNT_TRY(entry_phase(ctx, ak, na, a_dims, bk, nb, b_dims, (strategy & LNT_SPGEMM_CSC) != 0, &E));
  LNT_TRY(ensure(ctx->aux, E.scan_words * 8 + 64));
  LNT_TRY(exclusive_scan<uint32_t>(ctx, E.a_len, E.prod_off, nb, E.prod_off + nb, ctx->aux, 0));
  LNT_TRY(exclusive_scan<uint32_t>(ctx, E.head, E.seg_idx, nb, E.seg_idx + nb, ctx->aux, 0));
  uint64_t hv[2];
  LNT_TRY(read_two_u64(ctx, E.prod_off + nb, E.seg_idx + nb, hv));
  const uint64_t nprod = hv[0], nseg = hv[1];
  if (nprod_host) *nprod_host = nprod;
  if (nprod >= (uint64_t(1) << 32)) {
    prof_end(ctx, slot);
    return fail(LNT_ERR_TOO_LARGE, "more than 2^32 scalar products in one lattice product");
  }
  if (nprod == 0) {
    prof_end(ctx, slot);
    return LNT_OK;
  }
  // The dense accumulator takes every multi-entry segment when its index
  // space fits in LDS: the rows (h <= kDenseRows), or for mtt_times the ranks
  // of A's rows among the rows present.  Otherwise those strategies use the
  // sort accumulator.  The expand-sort-compress strategy (outer_times) sends
  // every multi-entry segment through the global sort.
  bool dense = false;
  RankTables rt{};
  uint64_t nrank = 0;
  if (acc == LNT_SPGEMM_MTT && ah <= (uint64_t(1) << 31)) {
    LNT_TRY(rank_count(ctx, ak, na, ah, E.scan_words, &rt, &nrank));
    if (nrank <= kDenseRows) {
      LNT_TRY(rank_tables(ctx, ak, na, ah, rt));
      dense = true;
    }
  } else if (acc == LNT_SPGEMM_DENSE && ah <= kDenseRows) {
    dense = true;
  }
  const uint32_t small_max = dense ? 0xFFFFFFFFu
                             : acc == LNT_SPGEMM_ESC ? 0u
                                                     : (uint32_t)std::min(ctx->tuning.spgemm_small, kSmall);
  // Every long segment has more than small_max products: a bound on their
  // number the host knows now (none when every multi-entry segment is short).
  const uint64_t n3_ub =
      small_max == 0xFFFFFFFFu ? 0 : std::min<uint64_t>(nseg, nprod / ((uint64_t)small_max + 1) + 1);
  seg_first_kernel<<<grid_for(nb), 256, 0, ctx->stream>>>(E.head, E.seg_idx, nb, E.seg_first);
  LNT_TRY(launch_fill_u32(ctx, E.counts, 0, 4));
  // E.head (free after seg_first_kernel) flags the large segments; their
  // list is built in ascending order, so that their pieces run in output
  // order; E.ex_off holds its positions until the global sort reuses it.
  // Single-entry segments that are long on average are copied by
  // single_copy_kernel (A shared through each XCD's L2), not by the output
  // kernel.
  const bool single_copy = ctx->tuning.spgemm_single_copy && nprod >= 64 * nseg;
  classify_kernel<<<grid_for(nseg), 256, 0, ctx->stream>>>(E.seg_first, nseg, nb, E.prod_off, small_max,
                                                            E.seg_u, E.list2, E.head, E.direct, E.counts,
                                                            single_copy ? 1 : 0);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint32_t>(ctx, E.head, E.ex_off, nseg, E.ex_off + nseg, ctx->aux, 0));
  large_list_kernel<<<grid_for(nseg), 256, 0, ctx->stream>>>(E.head, E.ex_off, nseg, E.list3);
  LNT_HIP(hipGetLastError());
  // Row pieces (every accumulator but ESC): planned before the wait below,
  // over at most nseg large segments, so that one wait also returns how
  // many segments the pieces cannot take.
  const bool try_pieces = acc != LNT_SPGEMM_ESC && ctx->tuning.spgemm_pieces;
  PieceArrays PA{};
  if (try_pieces) {
    LNT_TRY(ensure(ctx->spg_plan, (size_t)nseg * (8 * 6 + 4 * 2) + 8 * 256));
    Carve c{(char *)ctx->spg_plan.ptr};
    PA.rmin = c.take<uint64_t>(nseg);
    PA.width = c.take<uint64_t>(nseg);
    PA.qc = c.take<uint64_t>(nseg);
    PA.qc_off = c.take<uint64_t>(nseg + 1);
    PA.ovf = c.take<uint32_t>(nseg);
    PA.list = c.take<uint32_t>(nseg);
    PA.seg_pos = c.take<uint64_t>(2 * nseg);
    piece_plan_kernel<<<grid_for(nseg * kWave), 256, 0, ctx->stream>>>(ak, E.a_lo, E.a_len, E.a_base, E.prod_off, E.seg_first,
                                                               nseg, nb, E.list3, E.ex_off + nseg, n3_ub, PA.rmin,
                                                               PA.width, PA.qc, PA.ovf, E.counts + 3, PA.seg_pos);
    LNT_HIP(hipGetLastError());
  }
  // With row pieces the host goes on without the lists' sizes: the kernels
  // below are sized by bounds and read them on the device (n2 in counts[0],
  // n3 in ex_off[nseg]); the fallback count comes back with nnz(C).
  const bool ahead = try_pieces && !wait_for_lists;
  uint32_t n2, n3, n_back = 0;
  if (ahead) {
    n2 = (uint32_t)std::min<uint64_t>(nseg, kSegGrid);
    n3 = (uint32_t)n3_ub;
  } else {
    LNT_HIP(hipMemcpyAsync(ctx->pinned + 4, E.counts, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipMemcpyAsync(ctx->pinned + 6, E.counts + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipMemcpyAsync(ctx->pinned + 8, E.ex_off + nseg, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    LNT_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t n3_64;
    memcpy(&n3_64, ctx->pinned + 8, sizeof(n3_64));
    n2 = std::min<uint32_t>(ctx->pinned[4], kSegGrid);
    n3 = (uint32_t)n3_64;
    n_back = ctx->pinned[6];
  }
  // n2: the grid of the short-segment kernels (they stride over the list);
  // n3: the exact number of long segments, or its bound (qc is zero past
  // the exact number, so scans and searches over n3 entries agree).
  if (n2 || n3) LNT_TRY(ensure(ctx->spg_tmp, nprod * 16));
  uint64_t *tmp_row = (uint64_t *)ctx->spg_tmp.ptr;
  double *tmp_val = tmp_row ? (double *)(tmp_row + nprod) : nullptr;
  if (n2 && dense) {
    if (acc == LNT_SPGEMM_MTT)
      dense_kernel<true><<<n2, kBlock, 0, ctx->stream>>>(ak, av, bv, E.a_lo, E.a_len, E.a_base, E.prod_off,
                                                         E.seg_first, nseg, nb, E.list2, E.counts, (uint32_t)nrank,
                                                         rt.ar, rt.rows, E.seg_u, tmp_row, tmp_val);
    else
      dense_kernel<false><<<n2, kBlock, 0, ctx->stream>>>(ak, av, bv, E.a_lo, E.a_len, E.a_base, E.prod_off,
                                                          E.seg_first, nseg, nb, E.list2, E.counts, (uint32_t)ah,
                                                          nullptr, nullptr, E.seg_u, tmp_row, tmp_val);
    LNT_HIP(hipGetLastError());
  } else if (n2) {
    if (acc == LNT_SPGEMM_HASH)
      small_hash_kernel<<<n2, kBlock, 0, ctx->stream>>>(ak, av, bv, E.a_lo, E.a_base, E.prod_off, E.seg_first,
                                                          nseg, nb, E.list2, E.counts, E.seg_u, tmp_row, tmp_val);
    else
      small_sort_kernel<<<n2, kBlock, 0, ctx->stream>>>(ak, av, bv, E.a_lo, E.a_base, E.prod_off, E.seg_first,
                                                          nseg, nb, E.list2, E.counts, E.seg_u, tmp_row, tmp_val);
    LNT_HIP(hipGetLastError());
  }
  // The global sort over the large segments in `list_esc` (expand with
  // composite keys, one radix sort, run sums).
  auto run_esc = [&](const uint32_t *list_esc, uint32_t n_esc) -> int {
    // Rows replaced by their ranks among A's rows when that takes the
    // composite keys below a radix-pass boundary or to 32 bits.
    uint64_t width = ah;
    const uint32_t *arank = nullptr, *rrows = nullptr;
    if (ctx->tuning.row_ranks && nprod >= (uint64_t(1) << 20) && ah <= (uint64_t(1) << 31)) {
      if (!rt.bm) LNT_TRY(rank_count(ctx, ak, na, ah, E.scan_words, &rt, &nrank));
      auto bits = [](unsigned __int128 v) {
        uint32_t b = 0;
        while (b < 128 && ((unsigned __int128)1 << b) < v) ++b;
        return b;
      };
      const uint32_t old_bits = bits((unsigned __int128)n_esc * ah), new_bits = bits((unsigned __int128)n_esc * nrank);
      if ((new_bits + 7) / 8 < (old_bits + 7) / 8 || (new_bits <= 32 && old_bits > 32)) {
        LNT_TRY(rank_tables(ctx, ak, na, ah, rt));
        width = nrank;
        arank = rt.ar;
        rrows = rt.rows;
      }
    }
    // Composite keys r*width + row must stay below 2^63; otherwise the large
    // segments go through in groups that satisfy it.
    uint32_t group = n_esc;
    while (group > 1 && (unsigned __int128)group * width >= ((unsigned __int128)1 << 63)) group = (group + 1) / 2;
    for (uint32_t g0 = 0; g0 < n_esc; g0 += group) {
      uint32_t gn = std::min(group, n_esc - g0);
      const uint32_t *l3 = list_esc + g0;
      large_sizes_kernel<<<(gn + 255) / 256, 256, 0, ctx->stream>>>(l3, gn, E.seg_first, nseg, nb, E.prod_off, E.np3);
      LNT_HIP(hipGetLastError());
      LNT_TRY(exclusive_scan<uint64_t>(ctx, E.np3, E.ex_off, gn, E.ex_off + gn, ctx->aux, 0));
      uint64_t total3;
      LNT_TRY(read_u64(ctx, E.ex_off + gn, &total3, 1));
      size_t scan_esc = scan_workspace_words(total3);
      LNT_TRY(ensure(ctx->spg_big, total3 * (8 + 8 + 4 + 8) + 1024));
      Carve c{(char *)ctx->spg_big.ptr};
      uint64_t *key3 = c.take<uint64_t>(total3);
      double *val3 = c.take<double>(total3);
      uint32_t *heads = c.take<uint32_t>(total3);
      uint64_t *grank = c.take<uint64_t>(total3 + 1);
      LNT_TRY(ensure(ctx->aux, std::max(E.scan_words, scan_esc) * 8 + 64));
      large_expand_kernel<<<chunk_grid(total3), 256, 0, ctx->stream>>>(
          ak, av, bv, E.a_lo, E.a_base, E.prod_off, E.seg_first, nseg, nb, l3, gn, E.ex_off, total3, width, arank,
          key3, val3);
      LNT_HIP(hipGetLastError());
      {
        // The composite keys are built above, below gn * width by
        // construction: no key-range check (and no host wait for it).
        ValidationOff quiet(ctx);
        LNT_TRY(sort_device(ctx, key3, val3, (uint32_t)total3, (uint64_t)gn * width, nullptr));
      }
      run_heads_kernel<<<grid_for(total3), 256, 0, ctx->stream>>>(key3, total3, heads);
      LNT_HIP(hipGetLastError());
      LNT_TRY(exclusive_scan<uint32_t>(ctx, heads, grank, total3, grank + total3, ctx->aux, 0));
      // np3 is free once ex_off is scanned: it takes the per-segment bases.
      large_base_kernel<<<(gn + 255) / 256, 256, 0, ctx->stream>>>(l3, gn, E.ex_off, grank, E.seg_first, E.prod_off,
                                                                    E.np3);
      large_reduce_kernel<<<grid_for(total3), 256, 0, ctx->stream>>>(key3, val3, heads, grank, total3,
                                                                      make_divider(width), rrows, E.np3, tmp_row,
                                                                      tmp_val);
      large_counts_kernel<<<(gn + 255) / 256, 256, 0, ctx->stream>>>(l3, gn, E.ex_off, grank, total3, E.seg_u);
      LNT_HIP(hipGetLastError());
    }
    return LNT_OK;
  };
  // The large segments: row pieces, with the global sort for the segments
  // the pieces cannot take (run first, so that every count the pieces do not
  // produce is known before they write).  When the caller's arrays hold
  // nprod entries (nnz(C) <= nprod), the pieces write their rows straight
  // into C: a look-back over all pieces in output order adds their counts to
  // the offsets of the other segments.
  const bool pieces = n3 && try_pieces;
  const bool direct = pieces && ck && !ck_alloc && capacity >= nprod;
  if (n3 && !pieces) LNT_TRY(run_esc(E.list3, n3));
  if (pieces) {
    const uint64_t q_ub = nprod / kPieceTarget + n3 + 1, pairs_ub = q_ub * kPieceLists;
    LNT_TRY(ensure(ctx->spg_piece, (pairs_ub + q_ub) * 8 + q_ub * 4 + 1024));
    Carve c{(char *)ctx->spg_piece.ptr};
    PA.bounds = c.take<uint64_t>(pairs_ub);
    PA.status = c.take<uint64_t>(q_ub);
    PA.piece_seg = c.take<uint32_t>(q_ub);
    LNT_TRY(ensure(ctx->aux, std::max(E.scan_words, scan_workspace_words(n3 + 1)) * 8 + 64));
    if (n_back) {
      LNT_TRY(launch_fill_u32(ctx, E.counts + 3, 0, 1));
      piece_fallback_kernel<<<(n3 + 255) / 256, 256, 0, ctx->stream>>>(E.list3, n3, PA.ovf, PA.list, E.counts + 3);
      LNT_HIP(hipGetLastError());
      LNT_TRY(run_esc(PA.list, n_back));
    }
    LNT_TRY(exclusive_scan<uint64_t>(ctx, PA.qc, PA.qc_off, n3, PA.qc_off + n3, ctx->aux, 0));
    uint64_t *out_part = E.np3;  // free once the global sort is done
    if (direct) LNT_TRY(exclusive_scan<uint64_t>(ctx, E.seg_u, out_part, nseg, out_part + nseg, ctx->aux, 0));
    fill_words_kernel<<<grid_for(q_ub, 4096), 256, 0, ctx->stream>>>(PA.status, q_ub);
    piece_bounds_kernel<<<grid_for(pairs_ub, 8192), 256, 0, ctx->stream>>>(ak, E.a_lo, E.a_len, E.a_base,
                                                                           E.seg_first, nseg, nb, E.list3, n3,
                                                                           PA.qc_off, PA.rmin, PA.width, PA.bounds,
                                                                           PA.piece_seg);
    piece_kernel<<<(uint32_t)q_ub, kPieceBlock, 0, ctx->stream>>>(
        ak, av, bv, E.a_lo, E.a_len, E.a_base, E.list3, n3, PA.qc_off, PA.rmin, PA.width, PA.bounds, PA.piece_seg,
        PA.seg_pos, E.counts + 2, PA.status, E.seg_u, tmp_row, tmp_val, direct ? 1 : 0, out_part, E.colid, ah, ck,
        cv);
    LNT_HIP(hipGetLastError());
    if (direct) {
      piece_total_kernel<<<(n3 + 255) / 256, 256, 0, ctx->stream>>>(E.list3, n3, ahead ? E.ex_off + nseg : nullptr,
                                                                     PA.qc_off, PA.status, E.seg_u, E.direct);
      LNT_HIP(hipGetLastError());
    }
  }
  LNT_TRY(exclusive_scan<uint64_t>(ctx, E.seg_u, E.out_off, nseg, E.out_off + nseg, ctx->aux, 0));
  // Output arrays of nprod entries hold C whatever its size (nnz(C) <=
  // nprod): the output launches are queued at once, sized for nprod, and
  // read nnz(C) on the device; the host reads it after them.  Otherwise
  // nnz(C) sizes the allocation or is checked against the capacity first.
  const bool queue_ahead = ck && !ck_alloc && capacity >= nprod;
  uint64_t nc = nprod;
  if (!queue_ahead) {
    uint64_t hv2[2];
    LNT_TRY(read_two_u64(ctx, E.out_off + nseg, (const uint64_t *)(E.counts + 2), hv2));
    nc = hv2[0];
    if (ahead && (hv2[1] >> 32)) {  // counts[3]: segments the pieces could not take
      prof_end(ctx, slot);
      return kRunAgainWaiting;
    }
  }
  if (ck_alloc) {
    void *k = nullptr, *v = nullptr;
    size_t bytes = (nc ? nc : 1) * 8;
    if (hipMalloc(&k, bytes) != hipSuccess || hipMalloc(&v, bytes) != hipSuccess) {
      (void)hipGetLastError();
      if (k) (void)hipFree(k);
      prof_end(ctx, slot);
      return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of the lattice product failed");
    }
    ck = (uint64_t *)k;
    cv = (double *)v;
    *ck_alloc = ck;
    *cv_alloc = cv;
  } else if (nc > capacity) {
    prof_end(ctx, slot);
    return fail(LNT_ERR_INVALID_ARGUMENT, "capacity " + std::to_string(capacity) + " is below nnz(C) = " +
                                              std::to_string(nc));
  }
  // nc: nnz(C), or its bound nprod when the launches are queued ahead.
  const uint64_t nchunks = (nc + kChunk - 1) / kChunk;
  LNT_TRY(ensure(ctx->spg_chunk, (nchunks + 1) * sizeof(uint64_t)));
  uint64_t *cs = (uint64_t *)ctx->spg_chunk.ptr;
  if (single_copy)
    single_copy_kernel<<<8 * (uint32_t)std::min<uint64_t>((nseg + kWave - 1) / kWave, 4096), 256, 0, ctx->stream>>>(
        ak, av, na, bv, E.a_lo, E.a_len, E.a_base, E.colid, E.seg_first, nseg, nb, E.out_off, ah, ck, cv);
  // With the single-entry segments copied and the long ones written by
  // their pieces (none staged, none through the global sort), the output
  // kernel has only the short segments, and leaves when there are none.
  const bool only_short = single_copy && (!n3 || (pieces && direct)) && n_back == 0;
  const uint32_t *n_short = only_short ? E.counts : nullptr;
  chunk_starts_kernel<<<grid_for(nchunks, 4096), 256, 0, ctx->stream>>>(E.out_off, nseg, cs, n_short);
  write_kernel<<<chunk_grid(nc), 256, 0, ctx->stream>>>(ak, av, bv, E.a_lo, E.a_base, E.colid, E.prod_off, E.seg_first,
                                                       nseg, nb, E.out_off, tmp_row, tmp_val, ah, cs, E.direct,
                                                       n_short, ck, cv);
  const hipError_t launch = hipGetLastError();
  if (launch != hipSuccess) {
    prof_end(ctx, slot);
    if (ck_alloc) {  // the result allocated above is not handed back
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(ck);
      (void)hipFree(cv);
      *ck_alloc = nullptr;
      *cv_alloc = nullptr;
    }
    return fail(LNT_ERR_DEVICE, std::string("write_kernel launch: ") + hipGetErrorString(launch));
  }
  prof_end(ctx, slot);
  if (queue_ahead) {
    uint64_t hv2[2];
    LNT_TRY(read_two_u64(ctx, E.out_off + nseg, (const uint64_t *)(E.counts + 2), hv2));
    nc = hv2[0];
    // Fallback segments: C is written again by a run that waits for them.
    if (ahead && (hv2[1] >> 32)) return kRunAgainWaiting;
  }
  prof_set_bytes(ctx, slot, 16.0 * ((double)na + nb + nc));
  *nc_host = nc;
  return LNT_OK;
}

}  // namespace

int lattice_mult_device(lnt_context ctx, const uint64_t *ak, const double *av, uint64_t na,
                        const int64_t a_dims[3], const uint64_t *bk, const double *bv, uint64_t nb,
                        const int64_t b_dims[3], int strategy, uint64_t *ck, double *cv,
                        uint64_t capacity, uint64_t *nc_host, uint64_t *nprod_host,
                        uint64_t **ck_alloc, double **cv_alloc) {
  int rc = lattice_mult_impl(ctx, ak, av, na, a_dims, bk, bv, nb, b_dims, strategy, ck, cv, capacity, nc_host,
                             nprod_host, ck_alloc, cv_alloc, false);
  if (rc == kRunAgainWaiting)
    rc = lattice_mult_impl(ctx, ak, av, na, a_dims, bk, bv, nb, b_dims, strategy, ck, cv, capacity, nc_host,
                           nprod_host, ck_alloc, cv_alloc, true);
  return rc;
}

}  // namespace lnt
