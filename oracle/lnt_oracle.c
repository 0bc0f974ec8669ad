/*
 * lnt_oracle.c -- CPU restatement of LibMIA's sparse hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Nothing in the product library links or calls this file.
 *
 * Reference paths are relative to src/multi_index_array/ unless they start
 * with src/.  Each function states what pins it (tests/test_oracle.py):
 *   - generators: exact integer/IEEE arithmetic, pinned by committed vectors;
 *   - lnto_change_linidx: closed form sub2ind(ind2sub) on exhaustive small dims;
 *   - sorts: ascending order + (key,value) multiset, against numpy argsort;
 *   - merges: dictionary union/intersection with the reference's op rules;
 *   - lattice product / contraction: symbolic boolean product (index set) and
 *     dense FP64 einsum (values within 1e-12, isEqualFuzzy semantics).
 * Parity unpinned: the exact FP64 bits of contraction sums in the reference
 * (they depend on its unstable per-column radix order; this file restates that
 * order swap for swap, but without running the reference nothing pins it).
 */
#include "lnt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================= synthetic-input generators ======================= */
/* Not reference code: the reference seeds boost::mt19937 from time(0)
 * (LibMIAUtil.h:43-46), which is neither reproducible nor portable.  These are
 * counter-based and use only integer ops and exact IEEE ops, so the GPU
 * generator in the bench reproduces them bit for bit. */

uint64_t lnto_mix64(uint64_t z) {
  /* SplitMix64 finaliser (Steele, Lea, Flood 2014). */
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t lnto_rand_u64(uint64_t seed, uint64_t i) {
  return lnto_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

double lnto_uniform_pm1(uint64_t seed, uint64_t i) {
  /* 53 random bits -> m * 2^-52 - 1 in [-1, 1 - 2^-52]; every step is exact. */
  uint64_t m = lnto_rand_u64(seed, i) >> 11;
  return (double)m * 0x1p-52 - 1.0;
}

double lnto_uniform_01(uint64_t seed, uint64_t i) {
  return (double)(lnto_rand_u64(seed, i) >> 11) * 0x1p-53;
}

double lnto_normal(uint64_t seed, uint64_t i) {
  /* Box-Muller on two independent 53-bit uniforms; u1 in (0,1]. */
  double u1 = (double)((lnto_rand_u64(seed, 2 * i) >> 11) + 1) * 0x1p-53;
  double u2 = (double)(lnto_rand_u64(seed, 2 * i + 1) >> 11) * 0x1p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

uint64_t lnto_feistel(uint64_t x, int bits, uint64_t seed) {
  /* Balanced 4-round Feistel network on [0, 2^bits), bits even: a bijection
   * for any round function. */
  int half = bits / 2;
  uint64_t mask = half >= 32 ? 0xFFFFFFFFull : ((1ull << half) - 1);
  uint64_t l = (x >> half) & mask, r = x & mask;
  for (int round = 0; round < 4; ++round) {
    uint64_t f = lnto_mix64(r ^ (seed + 0xA24BAED4963EE407ull * (uint64_t)(round + 1))) & mask;
    uint64_t nl = r;
    r = l ^ f;
    l = nl;
  }
  return (l << half) | r;
}

uint64_t lnto_permute_index(uint64_t i, uint64_t domain, uint64_t seed) {
  /* Bijection on [0, domain) by cycle walking a Feistel permutation of the
   * next even power of two. */
  if (domain <= 1) return 0;
  int bits = 0;
  while (bits < 64 && (1ull << bits) < domain) ++bits;
  if (bits & 1) ++bits;
  if (bits < 2) bits = 2;
  uint64_t x = i;
  do {
    x = lnto_feistel(x, bits, seed);
  } while (x >= domain);
  return x;
}

void lnto_gen_distinct_keys(int64_t *keys, int64_t n, uint64_t domain,
                            uint64_t seed, uint64_t offset) {
  for (int64_t t = 0; t < n; ++t)
    keys[t] = (int64_t)lnto_permute_index(offset + (uint64_t)t, domain, seed);
}

void lnto_gen_uniform_vals(double *vals, int64_t n, uint64_t seed,
                           uint64_t offset) {
  for (int64_t t = 0; t < n; ++t)
    vals[t] = lnto_uniform_pm1(seed, offset + (uint64_t)t);
}

void lnto_random_perm(int32_t *perm, int n, uint64_t seed) {
  /* Fisher-Yates with an unbiased-enough multiply-high draw (n is tiny). */
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int i = n - 1; i > 0; --i) {
    uint64_t r = lnto_rand_u64(seed, (uint64_t)i);
    int j = (int)(((unsigned __int128)r * (uint64_t)(i + 1)) >> 64);
    int32_t t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
}

static uint64_t hash_slot(uint64_t key, uint64_t mask) {
  return lnto_mix64(key ^ 0x5851F42D4C957F2Dull) & mask;
}

int lnto_gen_coords(int64_t *keys, double *vals, int64_t nnz, int order,
                    const int64_t *dims, const int32_t *dist, double zipf_s,
                    int value_kind, uint64_t seed) {
  if (order <= 0 || order > 16 || nnz < 0) return -1;
  double total = 1.0;
  for (int m = 0; m < order; ++m) total *= (double)dims[m];
  if ((double)nnz > total) return -1;
  /* Zipf tables: P(x) proportional to (x+1)^-s over [0, dim). */
  double *cdf[16] = {0};
  for (int m = 0; m < order; ++m) {
    if (!dist[m]) continue;
    cdf[m] = (double *)malloc(sizeof(double) * (size_t)dims[m]);
    double acc = 0.0;
    for (int64_t x = 0; x < dims[m]; ++x) {
      acc += pow((double)(x + 1), -zipf_s);
      cdf[m][x] = acc;
    }
    for (int64_t x = 0; x < dims[m]; ++x) cdf[m][x] /= acc;
  }
  uint64_t cap = 1;
  while (cap < (uint64_t)(2 * nnz + 2)) cap <<= 1;
  uint64_t *table = (uint64_t *)malloc(sizeof(uint64_t) * cap);
  for (uint64_t s = 0; s < cap; ++s) table[s] = ~0ull;
  int64_t got = 0;
  uint64_t draw = 0, max_draws = 1000ull * (uint64_t)(nnz + 1);
  int rc = 0;
  while (got < nnz) {
    if (draw >= max_draws) { rc = -1; break; }
    int64_t key = 0, stride = 1;
    for (int m = 0; m < order; ++m) {
      uint64_t r = lnto_rand_u64(seed, draw * (uint64_t)order + (uint64_t)m);
      int64_t sub;
      if (!dist[m]) {
        sub = (int64_t)(((unsigned __int128)r * (uint64_t)dims[m]) >> 64);
      } else {
        double u = (double)(r >> 11) * 0x1p-53;
        int64_t lo = 0, hi = dims[m] - 1; /* first x with cdf[x] > u */
        while (lo < hi) {
          int64_t mid = (lo + hi) / 2;
          if (cdf[m][mid] > u) hi = mid; else lo = mid + 1;
        }
        sub = lo;
      }
      key += sub * stride;
      stride *= dims[m];
    }
    ++draw;
    uint64_t s = hash_slot((uint64_t)key, cap - 1);
    int dup = 0;
    while (table[s] != ~0ull) {
      if (table[s] == (uint64_t)key) { dup = 1; break; }
      s = (s + 1) & (cap - 1);
    }
    if (dup) continue;
    table[s] = (uint64_t)key;
    keys[got] = key;
    vals[got] = value_kind == 1 ? lnto_normal(seed ^ 0xD1B54A32D192ED03ull, (uint64_t)got)
                                : lnto_uniform_pm1(seed ^ 0xD1B54A32D192ED03ull, (uint64_t)got);
    ++got;
  }
  free(table);
  for (int m = 0; m < order; ++m) free(cdf[m]);
  return rc;
}

/* ================================ index math =============================== */

/* IndexUtil.h:866-879 sub2ind_reorder. */
int64_t lnto_sub2ind(const int64_t *sub, const int32_t *seq,
                     const int64_t *dims, int order) {
  uint64_t key = 0, mult = 1;
  for (int t = 0; t < order; ++t) {
    key += (uint64_t)sub[seq[t]] * mult;
    mult *= (uint64_t)dims[seq[t]];
  }
  return (int64_t)key;
}

/* IndexUtil.h:318-333 ind2sub_reorder. */
void lnto_ind2sub(int64_t key, const int32_t *seq, const int64_t *dims,
                  int order, int64_t *sub) {
  uint64_t k = (uint64_t)key;
  for (int t = 0; t < order; ++t) {
    sub[seq[t]] = (int64_t)(k % (uint64_t)dims[seq[t]]);
    k /= (uint64_t)dims[seq[t]];
  }
}

/* SparseMIABase.h:768-814 change_linIdx_sequence with the relinearisation of
 * IndexUtil.h:540-598 reShuffleLinearIndex: peel the old-order digits from the
 * most significant down (dividing by the old-order place values, the
 * dim_accumulator of create_shuffle_needs, IndexUtil.h:480-487) and re-weight
 * each by that mode's place value in the new order.  Does not sort. */
void lnto_change_linidx(int64_t *keys, int64_t n, int order,
                        const int64_t *dims, const int32_t *cur_seq,
                        const int32_t *new_seq) {
  if (order <= 1) return;
  int same = 1;
  for (int t = 0; t < order; ++t) same &= cur_seq[t] == new_seq[t];
  if (same) return;
  uint64_t old_place[64], new_weight[64];
  uint64_t acc = 1;
  for (int t = 0; t < order; ++t) {
    old_place[t] = acc;
    acc *= (uint64_t)dims[cur_seq[t]];
  }
  uint64_t stride_of_mode[64];
  acc = 1;
  for (int t = 0; t < order; ++t) {
    stride_of_mode[new_seq[t]] = acc;
    acc *= (uint64_t)dims[new_seq[t]];
  }
  for (int t = 0; t < order; ++t) new_weight[t] = stride_of_mode[cur_seq[t]];
  for (int64_t i = 0; i < n; ++i) {
    uint64_t idx = (uint64_t)keys[i], out = 0;
    for (int t = order - 1; t >= 1; --t) {
      uint64_t q = idx / old_place[t];
      out += q * new_weight[t];
      idx -= q * old_place[t];
    }
    out += idx * new_weight[0];
    keys[i] = (int64_t)out;
  }
}

/* ================================= sorting ================================ */
/* Introsort with a follower payload: LibMIAAlgorithm.h:153-174 (DualSwapper),
 * :439-511 (Partition), :522-535 (HeapSort/SiftDown), :547-560 (MedianOfThree),
 * :570-611 (IntrosortRec), :621-632 (IntrosortDepth).  Same comparisons and
 * swaps in the same order, so equal keys end up where the reference puts them. */

static inline void dual_swap(int64_t *k, double *v, int64_t i, int64_t j) {
  double tv = v[i];
  v[i] = v[j];
  v[j] = tv;
  int64_t tk = k[i];
  k[i] = k[j];
  k[j] = tk;
}

static void sift_down(int64_t *k, double *v, int64_t size, int64_t cur) {
  int64_t root = cur;
  while (root * 2 + 1 < size) {
    int64_t child = root * 2 + 1;
    if (child + 1 < size && k[child] < k[child + 1]) child += 1;
    if (k[root] < k[child]) {
      dual_swap(k, v, child, root);
      root = child;
    } else {
      return;
    }
  }
}

static void heap_sort(int64_t *k, double *v, int64_t size) {
  for (int64_t s = (size - 2) / 2; s >= 0; --s) sift_down(k, v, size, s);
  for (int64_t e = size - 1; e > 0; --e) {
    dual_swap(k, v, e, 0);
    sift_down(k, v, e, 0);
  }
}

static int64_t median_of_three(const int64_t *k, int64_t one, int64_t two,
                               int64_t three) {
  int c12 = k[one] < k[two], c13 = k[one] < k[three], c23 = k[two] < k[three];
  if (c12 && c23) return two;
  if (c12 && !c23 && c13) return three;
  if (!c12 && c13) return one;
  if (!c12 && !c13 && c23) return three;
  if (c12 && !c13) return one;
  return two;
}

static int64_t partition_pairs(int64_t *k, double *v, int64_t b, int64_t e) {
  int64_t lhs = b + 1, rhs = e - 1;
  for (;;) {
    while (lhs < rhs && !(k[rhs] < k[b])) --rhs;
    while (lhs < rhs && k[lhs] < k[b]) ++lhs;
    if (lhs == rhs) break;
    dual_swap(k, v, lhs, rhs);
  }
  if (!(k[lhs] < k[b])) return b;
  dual_swap(k, v, b, lhs);
  return lhs;
}

static uint64_t introsort_depth(int64_t n) {
  uint64_t lg2 = 0;
  for (uint64_t m = (uint64_t)n; m != 0; m >>= 1) ++lg2;
  return lg2 * 2;
}

static void introsort_rec(int64_t *k, double *v, int64_t b, int64_t e,
                          uint64_t depth) {
  const int64_t block = 50;
  for (;;) {
    int64_t n = e - b;
    if (n < block) return;
    if (depth == 0) {
      heap_sort(k + b, v + b, n);
      return;
    }
    int64_t pivot = median_of_three(k, b, b + n / 2, e - 1);
    dual_swap(k, v, pivot, b);
    int64_t p = partition_pairs(k, v, b, e);
    introsort_rec(k, v, b, p, depth - 1);
    b = p + 1; /* second recursive call of the reference, as a loop */
    depth -= 1;
  }
}

/* LibMIAAlgorithm.h:673-706 InsertionSortImproved (stable shifting insert). */
static void insertion_sort_improved(int64_t *k, double *v, int64_t n) {
  for (int64_t i = 1; i < n; ++i) {
    if (k[i] < k[i - 1]) {
      int64_t ck = k[i];
      double cv = v[i];
      k[i] = k[i - 1];
      v[i] = v[i - 1];
      int64_t j = i - 1;
      for (; j > 0 && ck < k[j - 1]; --j) {
        k[j] = k[j - 1];
        v[j] = v[j - 1];
      }
      k[j] = ck;
      v[j] = cv;
    }
  }
}

/* LibMIARadix.h:119-180: in-place MSD radix, 2048 bins (11-bit digits),
 * cycle-leader permutation with the follower swapped alongside; bins of at
 * least 3000 recurse, bins of 50..2999 go to IntrosortRec, smaller bins are left
 * for the final insertion sort.  The Threshold template argument of the
 * reference's call sites is the intended 3000 (see the arity defect at
 * LibMIARadix.h:470). */
enum { RADIX = 2048, LOG_RADIX = 11, RADIX_THRESHOLD = 3000, INSERT_THRESHOLD = 50 };

static void radix_l1(int64_t *k, double *v, int64_t len, unsigned shift) {
  int64_t *count = (int64_t *)calloc(RADIX, sizeof(int64_t));
  int64_t *start = (int64_t *)malloc(sizeof(int64_t) * (RADIX + 1));
  int64_t *end = (int64_t *)malloc(sizeof(int64_t) * (RADIX + 1));
  const uint64_t mask = RADIX - 1;
  for (int64_t i = 0; i < len; ++i) count[((uint64_t)k[i] >> shift) & mask]++;
  start[0] = end[0] = 0;
  start[RADIX] = -1; /* sentinel */
  for (int b = 1; b < RADIX; ++b) start[b] = end[b] = start[b - 1] + count[b - 1];
  int64_t next_bin = 1;
  for (int64_t cur = 0; cur < len;) {
    uint64_t d;
    while (end[d = (((uint64_t)k[cur] >> shift) & mask)] != cur) {
      dual_swap(k, v, cur, end[d]);
      end[d]++;
    }
    end[d]++;
    while (end[next_bin - 1] == start[next_bin]) ++next_bin;
    cur = end[next_bin - 1];
  }
  if (shift != 0) {
    shift -= LOG_RADIX;
    for (int b = 0; b < RADIX; ++b) {
      int64_t nb = end[b] - start[b];
      if (nb >= RADIX_THRESHOLD)
        radix_l1(k + start[b], v + start[b], nb, shift);
      else if (nb >= INSERT_THRESHOLD)
        introsort_rec(k + start[b], v + start[b], 0, nb, introsort_depth(nb));
    }
  }
  free(count);
  free(start);
  free(end);
}

/* LibMIARadix.h:444-484 RadixSortInPlace_PowerOf2Radix_Unsigned: the digit
 * plan comes from ceil(log2(max_size)) bits, the top digit first. */
void lnto_radix_sort_pairs(int64_t *keys, double *vals, int64_t n,
                           uint64_t max_size) {
  if (n <= 0) return;
  unsigned long bits = max_size <= 1 ? 0 : (unsigned long)ceil(log2((double)max_size));
  unsigned long shifts = (unsigned long)ceil((double)bits / LOG_RADIX);
  unsigned long shift = bits >= LOG_RADIX ? (shifts - 1) * LOG_RADIX : 0;
  if (n >= RADIX_THRESHOLD)
    radix_l1(keys, vals, n, (unsigned)shift);
  else
    introsort_rec(keys, vals, 0, n, introsort_depth(n));
  insertion_sort_improved(keys, vals, n);
}

/* LibMIAAlgorithm.h:752-764 Introsort with DualSwapper (keys are unique where
 * the reference calls it, so its final insertion pass is interchangeable with
 * the improved one). */
void lnto_introsort_pairs(int64_t *keys, double *vals, int64_t n) {
  if (n <= 0) return;
  introsort_rec(keys, vals, 0, n, introsort_depth(n));
  insertion_sort_improved(keys, vals, n);
}

static uint64_t dimensionality(const int64_t *dims, int order) {
  uint64_t d = 1;
  for (int m = 0; m < order; ++m) d *= (uint64_t)dims[m];
  return d;
}

/* SparseMIABase.h:590-645 sort(seq) = change_sort_order.  Unsorted input:
 * change_linIdx_sequence then the MSD radix (:596-599, :727-728).  Sorted input
 * of fewer than 3000 entries: relinearise then Introsort (:615-618).  Sorted
 * larger input: the reference runs RadixShuffle::permute (:622-634,
 * LibMIARadix.h:811-900); with unique keys its output is identical to the
 * relinearise + MSD radix path restated here, which is what this oracle times. */
/* ============================== RadixShuffle ============================== */
/* SparseMIABase::sort(seq) on a sorted tensor (SparseMIABase.h:605-634):
 * relinearise, then RadixShuffle::permute (LibMIARadix.h:705-927, find and sort
 * stages :1107-1674, cost model :611-665, 786-807) with the stages that
 * setupPermute (IndexUtil.h:647-734) derives from the shuffle sequence.  The
 * control flow is restated branch for branch: the runs estimates that choose
 * between the stage-wise shuffle and a straight radix sort, the find stages
 * (runs of equal leading digits, already grouped by the sorted input), the
 * sort stages (an out-of-place counting pass on the stage digit, taken with the
 * power-of-two approximation of the divisor and repaired by a rotation), the
 * eliminators and the ping-pong between the caller's arrays and the scratch.
 * Keys are unique, so whatever path runs, a correct sort gives the full
 * sort's result.  One defect, followed here as written: a sort stage takes its
 * digits as key >> floor(log2(divisor)) (the divisor approximated by a power
 * of two, :1402) but plans its passes from the stage's size alone (the
 * constructor's mShiftRightAmounts, :762-772); when the divisor is not a
 * power of two, size * divisor >> floor(log2(divisor)) can need one bit more
 * than those passes take, the top bit is masked off, and bins interleave
 * beyond the one rotation per bin that repairs the approximation (:1414-1425):
 * e.g. dims {19,25,35,3}, seq {2,3,0,1} -> {1,3,0,2}, 13,495 keys.  Power-of-two
 * dims make the approximation exact. */
enum { SH_MAX_STAGES = 64 };

typedef struct {
  int nst;
  uint64_t div[SH_MAX_STAGES], maxd[SH_MAX_STAGES];
  unsigned shift[SH_MAX_STAGES], dshift[SH_MAX_STAGES];
  uint64_t total_max;
  unsigned total_shift;
  int first_sort;
  int64_t *s1;
  double *s2;
  int64_t count[RADIX + 1];
} shuffle_t;

static unsigned radix_start_shift(uint64_t max_size) {
  unsigned bits = max_size <= 1 ? 0 : (unsigned)ceil(log2((double)max_size));
  unsigned shifts = (unsigned)ceil((double)bits / LOG_RADIX);
  return bits >= LOG_RADIX ? (shifts - 1) * LOG_RADIX : 0;
}

/* IndexUtil.h:647-734 setupPermute: stages from the most significant index
 * down; an index must be sorted when an earlier (less significant) new index
 * came from a later old position; consecutive indices of one kind merge;
 * stages of size 1 are dropped, and so is a trailing find stage. */
static int setup_permute(const int *rss, const uint64_t *dims, int order, shuffle_t *S) {
  uint64_t dl[LNTO_MAX_ORDER];
  int sf[LNTO_MAX_ORDER], kind[SH_MAX_STAGES];
  dl[0] = 1;
  for (int i = 1; i < order; ++i) dl[i] = dl[i - 1] * dims[i - 1];
  for (int i = order - 1; i > 0; --i) {
    sf[i] = 0;
    for (int j = 0; j < i; ++j)
      if (rss[j] > rss[i]) { sf[i] = 1; break; }
  }
  int n = 0, cur = order - 1;
  while (cur > 0) {
    const int want = sf[cur];
    if (!n || kind[n - 1] != want) {
      kind[n] = want;
      S->div[n] = dl[cur];
      S->maxd[n] = dims[cur];
      ++n;
    }
    int t = cur - 1;
    while (t > 0 && sf[t] == want) {
      S->div[n - 1] = dl[t];
      S->maxd[n - 1] *= dims[t];
      --t;
    }
    cur = t;
    if (S->maxd[n - 1] == 1) --n;
  }
  if (n && kind[n - 1] == 0) --n;
  S->nst = n;
  return n ? kind[0] : 0;
}

/* :611-641 estimateNumberOfRuns (length is updated, as the reference's
 * reference parameter is). */
static float est_runs(uint64_t *length, uint64_t max_size) {
  float runs = 0;
  int radix_runs = (int)ceil(log2((double)max_size) / log2((double)RADIX));
  float first_run = (float)max_size / (float)(uint64_t)pow((double)RADIX, radix_runs - 1);
  if (*length > RADIX_THRESHOLD) {
    *length = (uint64_t)(*length / first_run);
    radix_runs--;
    runs += 1;
    while (*length > RADIX_THRESHOLD && radix_runs > 0) {
      runs += 1;
      radix_runs--;
      *length /= RADIX;
    }
  }
  if (radix_runs > 0 && *length <= RADIX_THRESHOLD) {
    if (*length > INSERT_THRESHOLD) runs += 1.5f;
    else if (*length > 1) runs += .5f;
    *length = 0;
  }
  return runs;
}

/* :646-665 getLengthThreshold. */
static uint64_t length_threshold(uint64_t max_size) {
  uint64_t t = 50;
  if (max_size == 1) return t;
  int radix_runs = (int)ceil(log2((double)max_size) / log2((double)RADIX));
  uint64_t first_run = max_size / (uint64_t)pow((double)RADIX, radix_runs - 1);
  for (int i = 0; i < radix_runs - 2; ++i) t *= RADIX;
  return t * first_run;
}

/* :786-807 estimateNumberOfRunsShuffle. */
static float est_runs_shuffle(const shuffle_t *S, uint64_t len, int start) {
  float runs = 0;
  for (int i = start; i < S->nst; i += 2) {
    uint64_t cur_max = (S->maxd[i] * S->div[i]) >> S->dshift[i];
    runs += est_runs(&len, cur_max);
    if (!len) break;
    if (i + 1 < S->nst) {
      float t = ceilf((float)len / (float)S->maxd[i + 1]);
      len = (uint64_t)(t > 2.0f ? t : 2.0f);
      runs += .5f;
    }
  }
  return runs;
}

/* :929-958 getNextShifts. */
static void next_shifts(const shuffle_t *S, int st, unsigned *straight_shift, uint64_t *straight_elim,
                        uint64_t *next_elim, uint64_t *shuffle_thresh) {
  const uint64_t cur_div = S->div[st], next_max = S->maxd[st + 1], straight_max = next_max * S->div[st + 1];
  *straight_elim = 0;
  *next_elim = 0;
  if (!(next_max && !(next_max & (next_max - 1)))) *next_elim = cur_div;
  if (!(straight_max && !(straight_max & (straight_max - 1)))) *straight_elim = cur_div;
  *straight_shift = radix_start_shift(straight_max);
  *shuffle_thresh = length_threshold(next_max);
}

/* :1001-1018 InsertionSortWithMove and :1020-1097 its eliminator variants. */
static void sift_down_move(int64_t *k2, double *v2, int64_t at, int64_t ck, double cv) {
  k2[at] = k2[at - 1];
  v2[at] = v2[at - 1];
  --at;
  for (; at > 0 && ck < k2[at - 1]; --at) {
    k2[at] = k2[at - 1];
    v2[at] = v2[at - 1];
  }
  k2[at] = ck;
  v2[at] = cv;
}

static void insertion_move(int64_t *k, double *v, int64_t *k2, double *v2, int64_t n, int64_t elim, int dest_first) {
  if (!n) return;
  if (dest_first) {
    for (int64_t i = 0; i < n; ++i) k[i] += elim;
    insertion_sort_improved(k, v, n);
    return;
  }
  k[0] += elim;
  k2[0] = k[0];
  v2[0] = v[0];
  for (int64_t i = 1; i < n; ++i) {
    k[i] += elim;
    if (k[i] < k2[i - 1]) {
      sift_down_move(k2, v2, i, k[i], v[i]);
    } else {
      k2[i] = k[i];
      v2[i] = v[i];
    }
  }
}

static void intro_then(int64_t *k, double *v, int64_t n) {
  if (n > INSERT_THRESHOLD) introsort_rec(k, v, 0, n, introsort_depth(n));
}

static void shuffle_sort(shuffle_t *S, int64_t *k1, double *v1, int64_t *k2, double *v2, int64_t n, unsigned cur_shift,
                         int st, int in_is_dest);

/* :1107-1245 RadixShuffleFind (later stages, data in k1, the other buffer k2). */
static void shuffle_find(shuffle_t *S, int64_t *k1, double *v1, int64_t *k2, double *v2, int64_t n, int st,
                         int in_is_dest) {
  if (st >= S->nst - 1) return;
  const uint64_t cur_div = S->div[st];
  uint64_t cur_val = (uint64_t)k1[0] / cur_div, next_val = (cur_val + 1) * cur_div;
  int64_t off = 0, cnt = 1;
  unsigned straight_shift;
  uint64_t straight_elim, next_elim, thresh;
  next_shifts(S, st, &straight_shift, &straight_elim, &next_elim, &thresh);
  for (int64_t it = 1; it <= n; ++it) {
    if (it < n && (uint64_t)k1[it] < next_val) {
      ++cnt;
      continue;
    }
    int64_t *a = k1 + off, *b = k2 + off;
    double *av = v1 + off, *bv = v2 + off;
    if (cnt >= RADIX_THRESHOLD) {
      const int64_t elim = (int64_t)(cur_val * ((uint64_t)cnt > thresh ? next_elim : straight_elim));
      for (int64_t i = 0; i < cnt; ++i) a[i] -= elim;
      if ((uint64_t)cnt > thresh) {
        shuffle_sort(S, a, av, b, bv, cnt, S->shift[st + 1], st + 1, in_is_dest);
        for (int64_t i = 0; i < cnt; ++i) a[i] += elim;
      } else {
        radix_l1(a, av, cnt, straight_shift);
        insertion_move(a, av, b, bv, cnt, elim, !in_is_dest);
      }
    } else {
      intro_then(a, av, cnt);
      insertion_move(a, av, b, bv, cnt, 0, !in_is_dest);
    }
    if (it == n) break;
    off = it;
    cnt = 1;
    if ((uint64_t)k1[it] < next_val + cur_div) {
      cur_val += 1;
      next_val += cur_div;
    } else {
      cur_val = (uint64_t)k1[it] / cur_div;
      next_val = (cur_val + 1) * cur_div;
    }
  }
}

/* :1399-1433 performRadixExtractDigitRotate: counting pass on the digit taken
 * with the divisor approximated by its power of two (dshift); after the last
 * digit of a stage, each bin is one rotation away from sorted. */
static void extract_digit_rotate(shuffle_t *S, int64_t *k1, double *v1, int64_t *k2, double *v2, int64_t n,
                                 unsigned cur_shift, int st) {
  const unsigned total = cur_shift + S->dshift[st];
  int64_t *c = S->count;
  memset(c, 0, sizeof(S->count));
  for (int64_t i = 0; i < n; ++i) c[(((uint64_t)k1[i] >> total) & (RADIX - 1)) + 1]++;
  for (int b = 1; b < RADIX + 1; ++b) c[b] += c[b - 1];
  int64_t c2[RADIX + 1];
  memcpy(c2, c, sizeof(c2));
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t d = ((uint64_t)k1[i] >> total) & (RADIX - 1);
    k2[c2[d]] = k1[i];
    v2[c2[d]++] = v1[i];
  }
  if (!cur_shift) {
    const int64_t check = st + 1 < S->nst ? (int64_t)S->div[st + 1] : 1;
    for (int b = 0; b < RADIX; ++b) {
      for (int64_t i = c[b]; i < c[b + 1] - 1; ++i) {
        if (k2[i] > k2[i + 1] && k2[i] - k2[i + 1] >= check) {
          /* std::rotate(first, middle = i + 1, last) on both arrays */
          const int64_t lo = c[b], mid = i + 1, hi = c[b + 1];
          int64_t *tk = (int64_t *)malloc((size_t)(hi - lo) * sizeof(int64_t));
          double *tv = (double *)malloc((size_t)(hi - lo) * sizeof(double));
          memcpy(tk, k2 + mid, (size_t)(hi - mid) * sizeof(int64_t));
          memcpy(tk + (hi - mid), k2 + lo, (size_t)(mid - lo) * sizeof(int64_t));
          memcpy(tv, v2 + mid, (size_t)(hi - mid) * sizeof(double));
          memcpy(tv + (hi - mid), v2 + lo, (size_t)(mid - lo) * sizeof(double));
          memcpy(k2 + lo, tk, (size_t)(hi - lo) * sizeof(int64_t));
          memcpy(v2 + lo, tv, (size_t)(hi - lo) * sizeof(double));
          free(tk);
          free(tv);
          break;
        }
      }
    }
  }
}

/* :1438-1530 RadixShuffleSort. */
static void shuffle_sort(shuffle_t *S, int64_t *k1, double *v1, int64_t *k2, double *v2, int64_t n, unsigned cur_shift,
                         int st, int in_is_dest) {
  int stage_done = 0;
  extract_digit_rotate(S, k1, v1, k2, v2, n, cur_shift, st);
  if (cur_shift == 0) {
    if (st + 2 >= S->nst) {
      if (!in_is_dest) {
        memcpy(k1, k2, (size_t)n * sizeof(int64_t));
        memcpy(v1, v2, (size_t)n * sizeof(double));
      }
      return;
    }
    st++;
    stage_done = 1;
  } else {
    cur_shift -= LOG_RADIX;
  }
  int64_t count[RADIX + 1];
  memcpy(count, S->count, sizeof(count));
  in_is_dest = !in_is_dest;
  for (int b = 0; b < RADIX; ++b) {
    const int64_t off = count[b], m = count[b + 1] - count[b];
    if (!m) continue;
    int64_t *a = k2 + off, *d = k1 + off;
    double *av = v2 + off, *dv = v1 + off;
    if (stage_done ? m >= INSERT_THRESHOLD : m >= RADIX_THRESHOLD) {
      if (stage_done) shuffle_find(S, a, av, d, dv, m, st, in_is_dest);
      else shuffle_sort(S, a, av, d, dv, m, cur_shift, st, in_is_dest);
    } else {
      intro_then(a, av, m);
      insertion_move(a, av, d, dv, m, 0, !in_is_dest);
    }
  }
}

/* :1247-1380 RadixShuffleFind, the first (in-place) call. */
static void shuffle_find_first(shuffle_t *S, int64_t *k, double *v, int64_t n) {
  const uint64_t cur_div = S->div[0];
  uint64_t cur_val = (uint64_t)k[0] / cur_div, next_val = (cur_val + 1) * cur_div;
  int64_t off = 0, cnt = 1;
  unsigned straight_shift;
  uint64_t straight_elim, next_elim, thresh;
  next_shifts(S, 0, &straight_shift, &straight_elim, &next_elim, &thresh);
  const uint64_t est_length = (uint64_t)n / S->maxd[0];
  const int num_runs = (int)est_runs_shuffle(S, est_length, 1);
  uint64_t t_length = est_length;
  const int num_straight = (int)est_runs(&t_length, S->total_max / S->maxd[0]);
  const int shuffle = num_runs < num_straight;
  for (int64_t it = 1; it <= n; ++it) {
    if (it < n && (uint64_t)k[it] < next_val) {
      ++cnt;
      continue;
    }
    int64_t *a = k + off;
    double *av = v + off;
    if (cnt >= RADIX_THRESHOLD) {
      const int64_t elim = (int64_t)(cur_val * (shuffle ? next_elim : straight_elim));
      for (int64_t i = 0; i < cnt; ++i) a[i] -= elim;
      if (shuffle) {
        shuffle_sort(S, a, av, S->s1, S->s2, cnt, S->shift[1], 1, 0);
        for (int64_t i = 0; i < cnt; ++i) a[i] += elim;
      } else {
        radix_l1(a, av, cnt, straight_shift);
        insertion_move(a, av, NULL, NULL, cnt, elim, 1);
      }
    } else {
      intro_then(a, av, cnt);
      if (cnt > 1) insertion_sort_improved(a, av, cnt);
    }
    if (it == n) break;
    off = it;
    cnt = 1;
    cur_val = (uint64_t)k[it] / cur_div;
    next_val = (cur_val + 1) * cur_div;
  }
}

/* :811-900 permute. */
static void shuffle_permute(shuffle_t *S, int64_t *k, double *v, int64_t n) {
  if (n < RADIX_THRESHOLD) {
    lnto_introsort_pairs(k, v, n);
    return;
  }
  uint64_t t_length = (uint64_t)n;
  const float num_straight = est_runs(&t_length, S->total_max);
  if (!S->first_sort) {
    if (S->nst <= 1) return;
    const uint64_t est_length = (uint64_t)n / S->maxd[0];
    if (est_length < INSERT_THRESHOLD) {
      radix_l1(k, v, n, S->total_shift);
      insertion_sort_improved(k, v, n);
    } else {
      shuffle_find_first(S, k, v, n);
    }
  } else if (est_runs_shuffle(S, (uint64_t)n, 0) < num_straight) {
    shuffle_sort(S, k, v, S->s1, S->s2, n, S->shift[0], 0, 0);
  } else {
    radix_l1(k, v, n, S->total_shift);
    insertion_sort_improved(k, v, n);
  }
}

/* Sorts keys already relinearised from cur_seq (in whose order they were
 * sorted) into new_seq: the RadixShuffle of SparseMIABase.h:622-634. */
static void radix_shuffle(int64_t *keys, double *vals, int64_t n, int order, const int64_t *dims,
                          const int32_t *cur_seq, const int32_t *new_seq) {
  shuffle_t *S = (shuffle_t *)calloc(1, sizeof(shuffle_t));
  int rss[LNTO_MAX_ORDER];
  uint64_t sdims[LNTO_MAX_ORDER];
  /* getShuffleSequence(old, new): for each new position, the old position of
   * its mode; reorder_from(dims, new_seq): dims in the new order. */
  for (int i = 0; i < order; ++i) {
    for (int j = 0; j < order; ++j)
      if (cur_seq[j] == new_seq[i]) rss[i] = j;
    sdims[i] = (uint64_t)dims[new_seq[i]];
  }
  S->first_sort = setup_permute(rss, sdims, order, S);
  S->total_max = dimensionality(dims, order);
  S->total_shift = radix_start_shift(S->total_max);
  for (int i = 0; i < S->nst; ++i) {
    S->shift[i] = radix_start_shift(S->maxd[i]);
    S->dshift[i] = (unsigned)floor(log2((double)S->div[i]));
  }
  S->s1 = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  S->s2 = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  shuffle_permute(S, keys, vals, n);
  free(S->s1);
  free(S->s2);
  free(S);
}

void lnto_sort_seq(int64_t *keys, double *vals, int64_t n, int order,
                   const int64_t *dims, const int32_t *cur_seq,
                   const int32_t *new_seq, int is_sorted) {
  int same = 1;
  for (int t = 0; t < order; ++t) same &= cur_seq[t] == new_seq[t];
  if (!is_sorted) {
    lnto_change_linidx(keys, n, order, dims, cur_seq, new_seq);
    lnto_radix_sort_pairs(keys, vals, n, dimensionality(dims, order));
    return;
  }
  if (order == 1 || same) return;
  lnto_change_linidx(keys, n, order, dims, cur_seq, new_seq);
  if (n < 3000)
    lnto_introsort_pairs(keys, vals, n);
  else
    lnto_radix_sort_pairs(keys, vals, n, dimensionality(dims, order));
}

/* The reference's own path of sort(seq) on sorted input of 3000 or more keys:
 * RadixShuffle (above).  Its output equals lnto_sort_seq's for power-of-two
 * dims; for others RadixShuffle's digits can overflow its passes (the
 * defect noted above radix_shuffle) and leave the keys unsorted, which is
 * why the checker's lnto_sort_seq sorts fully. */
void lnto_sort_seq_shuffle(int64_t *keys, double *vals, int64_t n, int order, const int64_t *dims,
                           const int32_t *cur_seq, const int32_t *new_seq, int is_sorted) {
  int same = 1;
  for (int t = 0; t < order; ++t) same &= cur_seq[t] == new_seq[t];
  if (!is_sorted || n < 3000) {
    lnto_sort_seq(keys, vals, n, order, dims, cur_seq, new_seq, is_sorted);
    return;
  }
  if (order == 1 || same) return;
  lnto_change_linidx(keys, n, order, dims, cur_seq, new_seq);
  radix_shuffle(keys, vals, n, order, dims, cur_seq, new_seq);
}

/* FunctionUtil.h:700-728 collect_duplicates_function with std::plus. */
int64_t lnto_collect_duplicates(int64_t *keys, double *vals, int64_t n) {
  if (n <= 0) return 0;
  int64_t r = 0;
  for (int64_t f = 1; f < n; ++f) {
    if (keys[r] != keys[f]) {
      ++r;
      keys[r] = keys[f];
      vals[r] = vals[f];
    } else {
      vals[r] = vals[r] + vals[f];
    }
  }
  return r + 1;
}

int64_t lnto_collect_duplicates_member(int64_t *keys, double *vals, int64_t n) {
  if (n <= 0) return 0;
  int64_t r = 0;
  for (int64_t f = 1; f < n; ++f) {
    if (keys[r] != keys[f]) {
      if (vals[f] != 0.0) {  /* `if(this->data_at(first))` */
        ++r;
        keys[r] = keys[f];
        vals[r] = vals[f];
      }
    } else {
      vals[r] = vals[r] + vals[f];
    }
  }
  return r + 1;
}

int64_t lnto_sort_merge(const int64_t *ak, const double *av, int64_t na, const int64_t *bk,
                        const double *bv, int64_t nb, int op, uint64_t bound, int64_t *ck,
                        double *cv) {
  for (int64_t i = 0; i < na; ++i) {
    ck[i] = ak[i];
    cv[i] = av[i];
  }
  for (int64_t i = 0; i < nb; ++i) {
    ck[na + i] = bk[i];
    cv[na + i] = op == LNTO_OP_SUB ? -1 * bv[i] : bv[i];
  }
  lnto_radix_sort_pairs(ck, cv, na + nb, bound);
  return lnto_collect_duplicates_member(ck, cv, na + nb);
}

/* ================================= merging ================================ */

static inline double merge_op(int op, double a, double b) {
  return op == LNTO_OP_SUB ? a - b : a + b;
}

/* FunctionUtil.h:373-419 outside_merge_sparse_storage_containers (op 0/1):
 * sorted union; equal keys -> op(a,b); A only -> a; B only -> op(0,b).
 * op 2: the pure Hadamard product of SparseMIABase.h:1717-1826 (no outer
 * indices): sorted intersection with a*b. */
int64_t lnto_merge(const int64_t *ak, const double *av, int64_t na,
                   const int64_t *bk, const double *bv, int64_t nb, int op,
                   int64_t *ck, double *cv) {
  int64_t i = 0, j = 0, c = 0;
  if (op == LNTO_OP_HADAMARD)
    return lnto_outer_inter_product(ak, av, na, 1, bk, bv, nb, 1, ck, cv);
  while (i < na && j < nb) {
    if (ak[i] < bk[j]) {
      ck[c] = ak[i]; cv[c++] = av[i++];
    } else if (bk[j] < ak[i]) {
      ck[c] = bk[j]; cv[c++] = merge_op(op, 0.0, bv[j++]);
    } else {
      ck[c] = ak[i]; cv[c++] = merge_op(op, av[i++], bv[j++]);
    }
  }
  while (j < nb) { ck[c] = bk[j]; cv[c++] = merge_op(op, 0.0, bv[j++]); }
  while (i < na) { ck[c] = ak[i]; cv[c++] = av[i++]; }
  return c;
}

/* src/MATLAB/mex/sparse_merge.cpp:16-85 (SparseMergeMex): the general merge
 * (latA + latB / latA - latB, :58-84) behind the mex's own shortcuts for an
 * empty operand: both empty -> empty (:26-28), B empty -> a copy of A
 * (:29-39), A empty -> B for op 0, -1*b for op 1 (:40-57). */
int64_t lnto_mex_merge(const int64_t *ak, const double *av, int64_t na,
                       const int64_t *bk, const double *bv, int64_t nb, int op,
                       int64_t *ck, double *cv) {
  if (na == 0 && nb == 0) return 0;
  if (nb == 0) {
    memcpy(ck, ak, (size_t)na * sizeof(*ck));
    memcpy(cv, av, (size_t)na * sizeof(*cv));
    return na;
  }
  if (na == 0) {
    for (int64_t j = 0; j < nb; ++j) {
      ck[j] = bk[j];
      cv[j] = op == 0 ? bv[j] : -1 * bv[j];
    }
    return nb;
  }
  return lnto_merge(ak, av, na, bk, bv, nb, op, ck, cv);
}

/* SparseMIABase.h:1717-1826 perform_noLatticeMult after both operands were
 * sorted to [outer, inter] order (:1746-1747): for every inter index present in
 * both, the outer product of the two groups, B outer loop, A inner loop, key
 * l_outer + r_outer*L + inter*L*R (:1811-1820).  Scans instead of the
 * reference's interpolation searches (find_start/end_idx, :2055-2132), which
 * only change how fast the group bounds are found.  ck == NULL counts only. */
int64_t lnto_outer_inter_product(const int64_t *ak, const double *av,
                                 int64_t na, int64_t l_outer_size,
                                 const int64_t *bk, const double *bv,
                                 int64_t nb, int64_t r_outer_size,
                                 int64_t *ck, double *cv) {
  int64_t i = 0, j = 0, c = 0;
  const uint64_t L = (uint64_t)l_outer_size, R = (uint64_t)r_outer_size;
  while (i < na && j < nb) {
    uint64_t ta = (uint64_t)ak[i] / L, tb = (uint64_t)bk[j] / R;
    if (ta < tb) { ++i; continue; }
    if (tb < ta) { ++j; continue; }
    int64_t ie = i, je = j;
    while (ie < na && (uint64_t)ak[ie] / L == ta) ++ie;
    while (je < nb && (uint64_t)bk[je] / R == tb) ++je;
    uint64_t base = ta * L * R;
    for (int64_t q = j; q < je; ++q) {
      uint64_t ro = (uint64_t)bk[q] % R;
      for (int64_t p = i; p < ie; ++p) {
        if (ck) {
          ck[c] = (int64_t)((uint64_t)ak[p] % L + ro * L + base);
          cv[c] = av[p] * bv[q];
        }
        ++c;
      }
    }
    i = ie;
    j = je;
  }
  return c;
}

int64_t lnto_outer_inter_count(const int64_t *ak, int64_t na,
                               int64_t l_outer_size, const int64_t *bk,
                               int64_t nb, int64_t r_outer_size) {
  return lnto_outer_inter_product(ak, NULL, na, l_outer_size, bk, NULL, nb,
                                  r_outer_size, NULL, NULL);
}

/* ============================ lattice product ============================= */
/* Lattice key = row + h*col + h*w*tab (SparseLatticeBase.h:867-877); operands
 * sorted ascending (column major).  nprod = sum over B entries (k,j,t) of
 * nnz(A(:,k,t)). */

static int64_t lower_bound_i64(const int64_t *a, int64_t n, int64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

int64_t lnto_lattice_nprod(const int64_t *ak, int64_t na, int64_t ah,
                           int64_t aw, const int64_t *bk, int64_t nb,
                           int64_t bh, int64_t bw) {
  int64_t total = 0;
  for (int64_t q = 0; q < nb; ++q) {
    int64_t t = bk[q] / (bh * bw);
    int64_t krow = bk[q] % bh;
    int64_t lo = lower_bound_i64(ak, na, t * ah * aw + krow * ah);
    int64_t hi = lower_bound_i64(ak, na, t * ah * aw + (krow + 1) * ah);
    total += hi - lo;
  }
  return total;
}

/* SparseLatticeBase.h:1854-1990 csc_no_accum<true>: for each tab present in
 * both lattices and each column j of B (in key order), scatter beta*A(:,k) for
 * every B entry (k,j) (mult_scatter, :1993-2031, value beta*a), then
 * RadixSortInPlace the column's rows with max_size = height (:1966-1969),
 * collect duplicates with plus (:1970-1972) and add tab*h*w + h*j (:1974-1978).
 * Returns nnz(C), or -1 if cap is too small. */
int64_t lnto_lattice_mult(const int64_t *ak, const double *av, int64_t na,
                          int64_t ah, int64_t aw, int64_t ad,
                          const int64_t *bk, const double *bv, int64_t nb,
                          int64_t bh, int64_t bw, int64_t bd, int64_t *ck,
                          double *cv, int64_t cap) {
  (void)ad; (void)bd;
  if (aw != bh) return -1;
  const int64_t a_tab_size = ah * aw, b_tab_size = bh * bw;
  int64_t c = 0, q = 0;
  while (q < nb) {
    int64_t tab = bk[q] / b_tab_size;
    int64_t col = (bk[q] % b_tab_size) / bh;
    int64_t col_end = tab * b_tab_size + (col + 1) * bh;
    int64_t start = c;
    for (; q < nb && bk[q] < col_end; ++q) {
      int64_t krow = bk[q] % bh;
      int64_t lo = lower_bound_i64(ak, na, tab * a_tab_size + krow * ah);
      int64_t hi = lower_bound_i64(ak, na, tab * a_tab_size + (krow + 1) * ah);
      if (c + (hi - lo) > cap) return -1;
      for (int64_t p = lo; p < hi; ++p) {
        ck[c] = ak[p] % ah;
        cv[c] = bv[q] * av[p];
        ++c;
      }
    }
    if (c - start > 1) {
      lnto_radix_sort_pairs(ck + start, cv + start, c - start, (uint64_t)ah);
      c = start + lnto_collect_duplicates(ck + start, cv + start, c - start);
    }
    int64_t off = tab * ah * bw + ah * col;
    for (int64_t p = start; p < c; ++p) ck[p] += off;
  }
  return c;
}

/* SparseLatticeBase.h:2224-2351 csc_times (and its scatter, mult_scatter
 * :2186-2220): a sparse accumulator (SPA) per tab -- row_marker[height],
 * reset per tab (:2283), and data_collector[height].  For each column j of B
 * (key order) and each B entry (k, j, beta) in key order, every row r of
 * A(:,k): if row_marker[r] < j+1 the row is new in this column (marker set,
 * row appended, collector[r] = beta*a), else collector[r] += beta*a.  The
 * appended rows are sorted (std::sort, :2322) and read out of the collector.
 * A row's sum therefore runs in ascending B row order, first product
 * unadded.  compress = 1 is mtt_times (:1579-1813): rows are first replaced
 * by their rank among the distinct rows of A's tab (a_row_map, :1697-1716) and
 * mapped back on output (:1800); the sums are the same operations.  The
 * reference's mtt_times re-purposes inner_index_map as its row marker with a
 * resize that keeps stale inner-index values (:1770-1771) instead of zeros, so
 * it can drop products of early columns; this restatement zeroes the marker
 * per tab, the evident intent.  CSC and DCSC column lookups (MultHelper
 * <.,false/true>, :121-169) only change how A(:,k) is found: both appear here
 * as one binary search.  Returns nnz(C), or -1 (mismatch or cap too small). */
static int cmp_i64(const void *x, const void *y) {
  const int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
  return (a > b) - (a < b);
}

int64_t lnto_lattice_mult_spa(const int64_t *ak, const double *av, int64_t na, int64_t ah, int64_t aw,
                              const int64_t *bk, const double *bv, int64_t nb, int64_t bh, int64_t bw,
                              int compress, int64_t *ck, double *cv, int64_t cap) {
  if (aw != bh) return -1;
  const int64_t a_tab_size = ah * aw, b_tab_size = bh * bw;
  int64_t *marker = NULL, *rowmap = NULL, *rank = NULL;
  double *collector = NULL;
  int64_t c = 0, q = 0, rc = 0;
  int64_t space = ah;
  if (compress) {
    rowmap = (int64_t *)malloc((size_t)(na > 0 ? na : 1) * sizeof(int64_t));
    rank = (int64_t *)malloc((size_t)(na > 0 ? na : 1) * sizeof(int64_t));
    if (!rowmap || !rank) { rc = -1; goto out; }
  }
  while (q < nb) {
    const int64_t tab = bk[q] / b_tab_size;
    const int64_t q_end = lower_bound_i64(bk, nb, (tab + 1) * b_tab_size);
    const int64_t a0 = lower_bound_i64(ak, na, tab * a_tab_size), a1 = lower_bound_i64(ak, na, (tab + 1) * a_tab_size);
    int64_t nrows = 0;
    if (compress) {
      /* a_row_map: the distinct rows of the tab's A entries, ascending; rank[p]
       * = position of A entry p's row in it. */
      for (int64_t p = a0; p < a1; ++p) rowmap[p - a0] = ak[p] % ah;
      qsort(rowmap, (size_t)(a1 - a0), sizeof(int64_t), cmp_i64);
      for (int64_t p = 0; p < a1 - a0; ++p)
        if (p == 0 || rowmap[p] != rowmap[nrows - 1]) rowmap[nrows++] = rowmap[p];
      for (int64_t p = a0; p < a1; ++p) {
        int64_t lo = 0, hi = nrows, r = ak[p] % ah;
        while (lo < hi) { int64_t mid = (lo + hi) / 2; if (rowmap[mid] < r) lo = mid + 1; else hi = mid; }
        rank[p] = lo;
      }
      space = nrows > 0 ? nrows : 1;
    }
    free(marker);
    free(collector);
    marker = (int64_t *)calloc((size_t)space, sizeof(int64_t));
    collector = (double *)malloc((size_t)space * sizeof(double));
    if (!marker || !collector) { rc = -1; goto out; }
    while (q < q_end) {
      const int64_t col = (bk[q] % b_tab_size) / bh;
      const int64_t col_end = tab * b_tab_size + (col + 1) * bh;
      const int64_t start = c;
      for (; q < q_end && bk[q] < col_end; ++q) {
        const int64_t krow = bk[q] % bh;
        const int64_t lo = lower_bound_i64(ak + a0, a1 - a0, tab * a_tab_size + krow * ah) + a0;
        const int64_t hi = lower_bound_i64(ak + a0, a1 - a0, tab * a_tab_size + (krow + 1) * ah) + a0;
        for (int64_t p = lo; p < hi; ++p) {
          const int64_t r = compress ? rank[p] : ak[p] % ah;
          if (marker[r] < col + 1) {
            if (c >= cap) { rc = -1; goto out; }
            marker[r] = col + 1;
            ck[c++] = r;
            collector[r] = bv[q] * av[p];
          } else {
            collector[r] += bv[q] * av[p];
          }
        }
      }
      qsort(ck + start, (size_t)(c - start), sizeof(int64_t), cmp_i64);
      const int64_t off = tab * ah * bw + ah * col;
      for (int64_t p = start; p < c; ++p) {
        cv[p] = collector[ck[p]];
        ck[p] = (compress ? rowmap[ck[p]] : ck[p]) + off;
      }
    }
  }
  rc = c;
out:
  free(marker);
  free(collector);
  free(rowmap);
  free(rank);
  return rc;
}

/* SparseLatticeBase.h:2034-2183 outer_times: A sorted column major, B sorted
 * row major (:2062-2063; given here column major and re-keyed to row major,
 * key = col + row*bw + tab*bh*bw, then sorted -- the order b.sort(RowMajor)
 * produces).  Per tab, A's columns and B's rows are walked together; for every
 * k present in both, every B entry of row k (ascending column) times every A
 * entry of column k (ascending row) is appended with key row + col*h and value
 * a*b (:2136-2146); the tab's products are sorted by the in-place MSD radix
 * with max_size h*bw (:2160), duplicates summed by collect_duplicates
 * (:2162) and the tab offset added (:2165-2166).  The reference's column /
 * row skips dereference one entry past the tab before testing the bound
 * (:2113, :2116); here the bound is tested first, with the same outcome. */
int64_t lnto_lattice_mult_outer(const int64_t *ak, const double *av, int64_t na, int64_t ah, int64_t aw,
                                const int64_t *bk, const double *bv, int64_t nb, int64_t bh, int64_t bw,
                                int64_t *ck, double *cv, int64_t cap) {
  if (aw != bh) return -1;
  const int64_t a_tab_size = ah * aw, b_tab_size = bh * bw, c_tab_size = ah * bw;
  int64_t *rk = (int64_t *)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(int64_t));
  double *rv = (double *)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(double));
  if (!rk || !rv) { free(rk); free(rv); return -1; }
  for (int64_t q = 0; q < nb; ++q) {
    const int64_t tab = bk[q] / b_tab_size, row = bk[q] % bh, col = (bk[q] % b_tab_size) / bh;
    rk[q] = col + row * bw + tab * b_tab_size;
    rv[q] = bv[q];
  }
  int64_t top = 0;
  for (int64_t q = 0; q < nb; ++q) top = rk[q] > top ? rk[q] : top;
  if (nb > 1) lnto_radix_sort_pairs(rk, rv, nb, (uint64_t)top + 1);
  int64_t c = 0, i = 0, j = 0, rc = 0;
  while (i < na && j < nb) {
    const int64_t ta = ak[i] / a_tab_size, tb = rk[j] / b_tab_size;
    if (ta < tb) { i = lower_bound_i64(ak, na, tb * a_tab_size); continue; }
    if (tb < ta) { j = lower_bound_i64(rk, nb, ta * b_tab_size); continue; }
    const int64_t tab = ta;
    const int64_t i_end = lower_bound_i64(ak, na, (tab + 1) * a_tab_size);
    const int64_t j_end = lower_bound_i64(rk, nb, (tab + 1) * b_tab_size);
    const int64_t start = c;
    int64_t ai = i, bj = j;
    while (ai < i_end && bj < j_end) {
      const int64_t acol = (ak[ai] % a_tab_size) / ah, brow = (rk[bj] % b_tab_size) / bw;
      if (acol < brow) { ++ai; continue; }
      if (brow < acol) { ++bj; continue; }
      int64_t a_e = ai, b_e = bj;
      while (a_e < i_end && (ak[a_e] % a_tab_size) / ah == acol) ++a_e;
      while (b_e < j_end && (rk[b_e] % b_tab_size) / bw == brow) ++b_e;
      if (c + (a_e - ai) * (b_e - bj) > cap) { rc = -1; goto out; }
      for (int64_t q = bj; q < b_e; ++q) {
        const int64_t bcol = rk[q] % bw;
        for (int64_t p = ai; p < a_e; ++p) {
          ck[c] = ak[p] % ah + bcol * ah;
          cv[c++] = av[p] * rv[q];
        }
      }
      ai = a_e;
      bj = b_e;
    }
    if (c - start) {
      lnto_radix_sort_pairs(ck + start, cv + start, c - start, (uint64_t)c_tab_size);
      c = start + lnto_collect_duplicates(ck + start, cv + start, c - start);
      for (int64_t p = start; p < c; ++p) ck[p] += tab * c_tab_size;
    }
    i = i_end;
    j = j_end;
  }
  rc = c;
out:
  free(rk);
  free(rv);
  return rc;
}

/* Index bookkeeping of the lattice product, MIA_Expr.h:449-561
 * (solve_product_expr_helper + product_lattice_expr_helper): A's rows are its
 * outer modes, columns its inner (contracted) modes, tabs its inter modes, each
 * in A's mode order; B's rows are its inner modes in the order matched to A's,
 * columns its outer modes in B's order, tabs its inter modes matched to A's.
 * C's dims are [A-outer, B-outer, inter] (:493-501). */
int lnto_contract_plan(int order_a, const int64_t *dims_a, const char *labels_a,
                       int order_b, const int64_t *dims_b, const char *labels_b,
                       int32_t *seq_a, int32_t *seq_b, int64_t *lat_a,
                       int64_t *lat_b, int64_t *dims_c, int *order_c) {
  int a_out[32], a_in[32], a_int[32], b_out[32], b_in[32], b_int[32];
  int na_out = 0, na_in = 0, na_int = 0, nb_out = 0;
  if (order_a > 32 || order_b > 32) return -1;
  for (int m = 0; m < order_a; ++m) {
    int match = -1;
    for (int r = 0; r < order_b; ++r)
      if (labels_b[r] == labels_a[m]) match = r;
    for (int r = 0; r < m; ++r)
      if (labels_a[r] == labels_a[m]) return -1; /* repeated label */
    int inter = labels_a[m] >= 'A' && labels_a[m] <= 'Z';
    if (match < 0) {
      if (inter) return -1;
      a_out[na_out++] = m;
    } else if (inter) {
      a_int[na_int] = m;
      b_int[na_int++] = match;
    } else {
      a_in[na_in] = m;
      b_in[na_in++] = match;
    }
    if (match >= 0 && dims_b[match] != dims_a[m]) return -1;
  }
  for (int r = 0; r < order_b; ++r) {
    int match = -1;
    for (int m = 0; m < order_a; ++m)
      if (labels_a[m] == labels_b[r]) match = m;
    for (int s = 0; s < r; ++s)
      if (labels_b[s] == labels_b[r]) return -1;
    if (match < 0) {
      if (labels_b[r] >= 'A' && labels_b[r] <= 'Z') return -1;
      b_out[nb_out++] = r;
    }
  }
  int t = 0;
  int64_t h = 1, w = 1, d = 1;
  for (int x = 0; x < na_out; ++x) { seq_a[t++] = a_out[x]; h *= dims_a[a_out[x]]; }
  for (int x = 0; x < na_in; ++x) { seq_a[t++] = a_in[x]; w *= dims_a[a_in[x]]; }
  for (int x = 0; x < na_int; ++x) { seq_a[t++] = a_int[x]; d *= dims_a[a_int[x]]; }
  lat_a[0] = h; lat_a[1] = w; lat_a[2] = d;
  t = 0;
  int64_t bwid = 1;
  for (int x = 0; x < na_in; ++x) seq_b[t++] = b_in[x];
  for (int x = 0; x < nb_out; ++x) { seq_b[t++] = b_out[x]; bwid *= dims_b[b_out[x]]; }
  for (int x = 0; x < na_int; ++x) seq_b[t++] = b_int[x];
  lat_b[0] = w; lat_b[1] = bwid; lat_b[2] = d;
  t = 0;
  for (int x = 0; x < na_out; ++x) dims_c[t++] = dims_a[a_out[x]];
  for (int x = 0; x < nb_out; ++x) dims_c[t++] = dims_b[b_out[x]];
  for (int x = 0; x < na_int; ++x) dims_c[t++] = dims_a[a_int[x]];
  *order_c = t;
  return 0;
}

/* MIA_Expr.h:864-925 lattice-product operator*: unfold both operands
 * (lattice_permutation_delegator, :339-419 -> SparseMIABase::sort(seq) through
 * toStraightLattice / toLatticeSort, SparseMIABase.h:1457-1586), multiply
 * (SparseLatticeBase::operator* -> csc_no_accum) and refold
 * (SparseMIA(dims, SparseLattice&&), SparseMIA.h:362-372: C keeps the lattice
 * keys, which are C's default-seq keys). */
int64_t lnto_contract(int64_t *ak, double *av, int64_t na, int order_a,
                      const int64_t *dims_a, const int32_t *seq_a,
                      int a_sorted, const char *labels_a, int64_t *bk,
                      double *bv, int64_t nb, int order_b,
                      const int64_t *dims_b, const int32_t *seq_b,
                      int b_sorted, const char *labels_b, int64_t *ck,
                      double *cv, int64_t cap, int64_t *dims_c, int *order_c) {
  int32_t la[32], lb[32];
  int64_t lat_a[3], lat_b[3];
  if (lnto_contract_plan(order_a, dims_a, labels_a, order_b, dims_b, labels_b,
                         la, lb, lat_a, lat_b, dims_c, order_c) != 0)
    return -1;
  lnto_sort_seq(ak, av, na, order_a, dims_a, seq_a, la, a_sorted);
  lnto_sort_seq(bk, bv, nb, order_b, dims_b, seq_b, lb, b_sorted);
  return lnto_lattice_mult(ak, av, na, lat_a[0], lat_a[1], lat_a[2], bk, bv, nb,
                           lat_b[0], lat_b[1], lat_b[2], ck, cv, cap);
}

/* FunctionUtil.h:688-698 isEqualFuzzy. */
int lnto_is_equal_fuzzy(double a, double b, double prec) {
  if (fabs(a) < 1 || fabs(b) < 1) return fabs(a - b) <= prec;
  double m = fabs(a) < fabs(b) ? fabs(a) : fabs(b);
  return fabs(a - b) <= m * prec;
}

/* below_tolerance (SparseMIABase.h:1170-1175) with SparseTolerance<double>
 * (LibMIAUtil.h:137-150). */
static int below_tolerance(double v) { return fabs(v) <= 1.11e-16; }

/* ======================= sparse x dense, no contraction ==================== */
/* SparseMIABase.h:1829-1910 noLatticeMult with a dense B.  A: keys sorted with
 * key = l + L*t (its outer modes, then its inter modes).  B: dense, first mode
 * fastest; r_modes[] are B's outer modes, i_modes[] its inter modes in the
 * order of A's.  For every A entry with a non-zero value and every B outer
 * index r in order whose B value is non-zero (:1881-1896): C entry a*b at key
 * r + l*R + t*L*R. */
int64_t lnto_sparse_dense_mult(const int64_t *ak, const double *av, int64_t na, int64_t L, int order_b,
                               const int64_t *b_dims, const double *b, const int32_t *r_modes, int nr,
                               const int32_t *i_modes, int ni, int64_t *ck, double *cv) {
  int64_t place[LNTO_MAX_ORDER], R = 1, T = 1, out = 0;
  place[0] = 1;
  for (int m = 1; m < order_b; ++m) place[m] = place[m - 1] * b_dims[m - 1];
  for (int q = 0; q < nr; ++q) R *= b_dims[r_modes[q]];
  for (int q = 0; q < ni; ++q) T *= b_dims[i_modes[q]];
  (void)T;
  for (int64_t e = 0; e < na; ++e) {
    const int64_t t = ak[e] / L, l = ak[e] % L;
    int64_t boff = 0, x = t;  /* B offset of the inter index (ind2sub over the inter dims) */
    for (int q = 0; q < ni; ++q) {
      boff += (x % b_dims[i_modes[q]]) * place[i_modes[q]];
      x /= b_dims[i_modes[q]];
    }
    if (!av[e]) continue;
    for (int64_t r = 0; r < R; ++r) {
      int64_t off = boff, y = r;
      for (int q = 0; q < nr; ++q) {
        off += (y % b_dims[r_modes[q]]) * place[r_modes[q]];
        y /= b_dims[r_modes[q]];
      }
      if (b[off]) {
        ck[out] = r + l * R + t * L * R;
        cv[out++] = av[e] * b[off];
      }
    }
  }
  return out;
}

/* ==================== unary contraction / attraction ===================== */
/* SparseMIABase.h:1913-2053 contract_attract.  group[m] = 0: mode m stays;
 * g in 1..ngroups: mode m is in group g; kind[g-1] 0 = contraction, 1 =
 * attraction.  The reference's index arrays are built here as its expression
 * code builds them: contraction positions grouped by group (increasing g),
 * then attraction positions (increasing g), then the staying modes in mode
 * order (:1920-1925), each group's positions ascending.  The tensor is sorted
 * in that order (:1928), then walked group by group exactly as :1966-2047:
 * the attraction check on the group's first entry, the running sum from 0 of
 * the entries whose contraction indices agree, pushed when non-zero, at key
 * other + attract_idx * other_dimensionality.  The pushed entries are
 * returned sorted by key (the reference leaves them unsorted).  keys/vals are
 * sorted in place as a side effect, as the reference sorts `this`. */
static int cmp_pair_key(const void *x, const void *y) {
  const int64_t a = ((const int64_t *)x)[0], b = ((const int64_t *)y)[0];
  return a < b ? -1 : a > b;
}

int64_t lnto_contract_attract(int64_t *keys, double *vals, int64_t n, int order, const int64_t *dims,
                              const int32_t *seq, const int32_t *group, const int32_t *kind, int ngroups,
                              int64_t *ck, double *cv) {
  int32_t sort_order[LNTO_MAX_ORDER], part_len[LNTO_MAX_ORDER], part_kind[LNTO_MAX_ORDER];
  int np = 0, t = 0, ncon = 0, natt = 0, nparts_con = 0;
  for (int pass = 0; pass < 2; ++pass)  /* contractions first, then attractions */
    for (int g = 1; g <= ngroups; ++g) {
      if (kind[g - 1] != pass) continue;
      int len = 0;
      for (int m = 0; m < order; ++m)
        if (group[m] == g) { sort_order[t++] = m; ++len; }
      part_len[np] = len;
      part_kind[np++] = pass;
      if (pass == 0) { ncon += len; ++nparts_con; } else natt += len;
    }
  const int nother = order - ncon - natt;
  for (int m = 0; m < order; ++m)
    if (group[m] == 0) sort_order[t++] = m;
  /* dimensionalities (:1929-1944) and the other modes' (:1922-1924) */
  uint64_t C = 1, A = 1, O = 1, att_range[LNTO_MAX_ORDER];
  int idx = 0;
  for (int p = 0; p < np; ++p) {
    uint64_t d = (uint64_t)dims[sort_order[idx]], pw = 1;
    for (int q = 0; q < part_len[p]; ++q) pw *= d;
    if (part_kind[p] == 0) C *= pw;
    else { att_range[p - nparts_con] = d; A *= pw; }
    idx += part_len[p];
  }
  for (int q = ncon + natt; q < order; ++q) O *= (uint64_t)dims[sort_order[q]];
  (void)nother;
  lnto_sort_seq(keys, vals, n, order, dims, seq, sort_order, 0);
  int64_t out = 0, cur = 0;
  int64_t sub[LNTO_MAX_ORDER], ind[LNTO_MAX_ORDER];
  while (cur < n) {
    lnto_ind2sub(keys[cur], sort_order, dims, order, sub);
    for (int q = 0; q < order; ++q) ind[q] = sub[sort_order[q]];
    /* attraction check of the current entry (:1955-1973) */
    int ok = 1, start = ncon;
    uint64_t attract_idx = 0, mult = 1;
    for (int p = nparts_con; p < np && ok; ++p) {
      for (int q = start + 1; q < start + part_len[p]; ++q)
        if (ind[q] != ind[start]) { ok = 0; break; }
      attract_idx += (uint64_t)ind[start] * mult;
      mult *= att_range[p - nparts_con];
      start += part_len[p];
    }
    if (!ok) { ++cur; continue; }
    const uint64_t head = (uint64_t)keys[cur];
    double sum = 0;
    while (cur < n && head / C == (uint64_t)keys[cur] / C) {
      lnto_ind2sub(keys[cur], sort_order, dims, order, sub);
      for (int q = 0; q < order; ++q) ind[q] = sub[sort_order[q]];
      int add = 1, cs = 0;
      for (int p = 0; p < nparts_con && add; ++p) {
        for (int q = cs + 1; q < cs + part_len[p]; ++q)
          if (ind[q] != ind[cs]) { add = 0; break; }
        cs += part_len[p];
      }
      if (add) sum += vals[cur];
      ++cur;
    }
    if (sum) {
      ck[out] = (int64_t)(head / (C * A) + attract_idx * O);
      cv[out++] = sum;
    }
  }
  /* sort the (key, value) pairs by key */
  int64_t *tmp = (int64_t *)malloc((size_t)(out ? out : 1) * 2 * sizeof(int64_t));
  for (int64_t q = 0; q < out; ++q) { tmp[2 * q] = ck[q]; memcpy(&tmp[2 * q + 1], &cv[q], sizeof(double)); }
  qsort(tmp, (size_t)out, 2 * sizeof(int64_t), cmp_pair_key);
  for (int64_t q = 0; q < out; ++q) { ck[q] = tmp[2 * q]; memcpy(&cv[q], &tmp[2 * q + 1], sizeof(double)); }
  free(tmp);
  return out;
}

/* SparseMIABase.h:1245-1296: the two-pointer walk, restated step for step. */
int lnto_compare_sparse(const int64_t *ak, const double *av, int64_t na, const int64_t *bk, const double *bv,
                        int64_t nb, int mode, double prec) {
  int64_t i = 0, j = 0;
  while (i < na && j < nb) {
    if (below_tolerance(av[i])) {
      ++i;
    } else if (below_tolerance(bv[j])) {
      ++j;
    } else {
      if (ak[i] != bk[j]) return 0;
      if (mode == 1 ? !lnto_is_equal_fuzzy(av[i], bv[j], prec) : !(av[i] == bv[j])) return 0;
      ++i;
      ++j;
    }
  }
  while (i < na)
    if (!below_tolerance(av[i++])) return 0;
  while (j < nb)
    if (!below_tolerance(bv[j++])) return 0;
  return 1;
}

/* ------------------------------ cost model ------------------------------ */
enum {
  MF = LNTO_MLP_F, MH = LNTO_MLP_H, MS = LNTO_MLP_S,
  M_W1 = 0, M_B1 = M_W1 + MH * MF, M_W2 = M_B1 + MH, M_B2 = M_W2 + MH * MH,
  M_W3 = M_B2 + MH, M_B3 = M_W3 + MS * MH, M_NW = M_B3 + MS, M_MEAN = M_NW,
  M_ISTD = M_NW + MF, M_NP = M_NW + 2 * MF
};

int64_t lnto_mlp_param_count(void) { return M_NP; }
int64_t lnto_mlp_weight_count(void) { return M_NW; }

/* Activations of one sample: xs (standardised input), z1, z2, y. */
static void mlp_sample(const double *P, const double *x, double *xs, double *z1, double *z2, double *y) {
  double h1[MH], h2[MH];
  for (int k = 0; k < MF; ++k) xs[k] = (x[k] - P[M_MEAN + k]) * P[M_ISTD + k];
  for (int j = 0; j < MH; ++j) {
    double z = P[M_B1 + j];
    for (int k = 0; k < MF; ++k) z += P[M_W1 + j * MF + k] * xs[k];
    z1[j] = z;
    h1[j] = z > 0 ? z : 0;
  }
  for (int j = 0; j < MH; ++j) {
    double z = P[M_B2 + j];
    for (int i = 0; i < MH; ++i) z += P[M_W2 + j * MH + i] * h1[i];
    z2[j] = z;
    h2[j] = z > 0 ? z : 0;
  }
  for (int o = 0; o < MS; ++o) {
    double v = P[M_B3 + o];
    for (int j = 0; j < MH; ++j) v += P[M_W3 + o * MH + j] * h2[j];
    y[o] = v;
  }
}

void lnto_mlp_forward(const double *params, const double *x, double *y) {
  double xs[MF], z1[MH], z2[MH];
  mlp_sample(params, x, xs, z1, z2, y);
}

double lnto_mlp_loss_grad(const double *P, const double *X, const double *T, int64_t n, double *grad) {
  int64_t m = 0;
  for (int64_t i = 0; i < n * MS; ++i) m += T[i] == T[i];
  for (int p = 0; p < M_NW; ++p) grad[p] = 0;
  if (m == 0) return 0;
  const double scale = 1.0 / (double)m;
  double loss = 0;
  for (int64_t s = 0; s < n; ++s) {
    double xs[MF], z1[MH], z2[MH], y[MS], d3[MS], dz2[MH], dz1[MH];
    mlp_sample(P, X + s * MF, xs, z1, z2, y);
    for (int o = 0; o < MS; ++o) {
      const double t = T[s * MS + o];
      d3[o] = 0;
      if (t == t) {
        const double e = y[o] - t;
        loss += scale * e * e;
        d3[o] = 2 * scale * e;
      }
    }
    for (int j = 0; j < MH; ++j) {
      double g = 0;
      for (int o = 0; o < MS; ++o) g += P[M_W3 + o * MH + j] * d3[o];
      dz2[j] = z2[j] > 0 ? g : 0;
    }
    for (int i = 0; i < MH; ++i) {
      double g = 0;
      for (int j = 0; j < MH; ++j) g += P[M_W2 + j * MH + i] * dz2[j];
      dz1[i] = z1[i] > 0 ? g : 0;
    }
    for (int o = 0; o < MS; ++o) {
      for (int j = 0; j < MH; ++j) grad[M_W3 + o * MH + j] += d3[o] * (z2[j] > 0 ? z2[j] : 0);
      grad[M_B3 + o] += d3[o];
    }
    for (int j = 0; j < MH; ++j) {
      for (int i = 0; i < MH; ++i) grad[M_W2 + j * MH + i] += dz2[j] * (z1[i] > 0 ? z1[i] : 0);
      grad[M_B2 + j] += dz2[j];
    }
    for (int j = 0; j < MH; ++j) {
      for (int k = 0; k < MF; ++k) grad[M_W1 + j * MF + k] += dz1[j] * xs[k];
      grad[M_B1 + j] += dz1[j];
    }
  }
  return loss;
}
