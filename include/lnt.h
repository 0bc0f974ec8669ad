/*
 * lnt.h -- C-ABI of the MI355X (gfx950) sparse multi-index-array library.
 *
 * The drop-in boundary for LibNT/LibMIA's sparse hot path: relinearisation,
 * radix sort and change_sort_order of (int64 linear index, double value)
 * pairs, sparse merge (+, -, Hadamard) and the sparse lattice product.  Every
 * entry point replaces one reference routine, cited below (paths relative to
 * the reference's src/).  The reference's own FFI for this path is the set of
 * MATLAB mex functions in src/MATLAB/mex/; the in-process C++ call sites are in
 * src/multi_index_array/.
 *
 * Conventions
 *   - Keys are int64 linear indices, always >= 0; values are double.  A key is
 *     computed in a mode order "seq" (0-based mode numbers, seq[0] = fastest
 *     varying mode): key = sum_t sub[seq[t]] * prod_{u<t} dims[seq[u]].
 *   - dims, seq and other small descriptors are host pointers.  Key/value
 *     arrays are host or device pointers according to the `mem` argument
 *     (LNT_MEM_HOST / LNT_MEM_DEVICE).  Device pointers must be on the
 *     context's device.
 *   - Every function returns an lnt_status.  On error nothing is written to the
 *     caller's arrays and lnt_last_error() describes the failure.  No function
 *     throws or calls exit/longjmp.  The error string is per thread.
 *   - Work is issued on the context's stream.  Functions that hand a count back
 *     to the host (merge, lattice product) synchronise that stream before
 *     returning; the in-place sorts do not (synchronise with
 *     lnt_context_synchronize or the stream when results are needed on the
 *     host).  Use one context per host thread.
 */
#ifndef LNT_H
#define LNT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LNT_ABI_VERSION 2
#define LNT_MAX_ORDER 16

typedef enum lnt_status {
  LNT_OK = 0,
  LNT_ERR_INVALID_ARGUMENT = 1, /* malformed sizes, seq not a permutation, ... */
  LNT_ERR_OUT_OF_MEMORY = 2,
  LNT_ERR_DEVICE = 3,           /* a HIP runtime call failed */
  LNT_ERR_NOT_SORTED = 4,       /* an input declared sorted is not */
  LNT_ERR_KEY_RANGE = 5,        /* a key is negative or >= the dimensionality */
  LNT_ERR_TOO_LARGE = 6,        /* size beyond the documented limits */
  LNT_ERR_DIMENSION_MISMATCH = 7,
  LNT_ERR_NO_DEVICE = 8
} lnt_status;

typedef enum lnt_memory { LNT_MEM_HOST = 0, LNT_MEM_DEVICE = 1 } lnt_memory;

typedef enum lnt_merge_op {
  LNT_OP_ADD = 0,     /* std::plus  */
  LNT_OP_SUB = 1,     /* std::minus */
  LNT_OP_HADAMARD = 2 /* element-wise product over the common support */
} lnt_merge_op;

typedef struct lnt_context_s *lnt_context;

/* ------------------------------------------------------------- context ---- */
/* Creates a context on HIP device `device`, issuing work on `stream` (a
 * hipStream_t; NULL = the device's null stream).  The context owns scratch
 * memory, tuning parameters and the learned-dispatch weights.  No reference
 * counterpart (the reference is single-threaded CPU code). */
int lnt_context_create(lnt_context *ctx, int device, void *stream);
int lnt_context_destroy(lnt_context ctx);
int lnt_context_set_stream(lnt_context ctx, void *stream);
int lnt_context_synchronize(lnt_context ctx);
/* Releases cached scratch memory. */
int lnt_context_trim(lnt_context ctx);
/* Input validation level.  0: none.  1 (default): sorts check every key
 * against the dimensionality inside their first read of the keys, before any
 * output is written (one stream synchronisation).  2: additionally, merges
 * check that both inputs are strictly ascending (one extra read of the keys). */
int lnt_context_set_validation(lnt_context ctx, int level);
const char *lnt_last_error(void);
int lnt_abi_version(void);
/* Kernel timing with HIP events recorded on the context's stream around each
 * launch.  Turning profiling on or off clears the records.  lnt_context_profile
 * sums, for one kernel ("histogram", "onesweep", "partition", "merge",
 * "relin"), the launches recorded, their total duration and the algorithmic
 * bytes they move (sort pass: 32 B per pair; histogram: 8 B per key; merge:
 * 16 B per input and per output element; relinearisation: 16 B per key).  No
 * reference counterpart (the reference has only ad-hoc std::chrono prints,
 * multi_index_array/SparseLatticeBase.h:1885,1981-1982). */
/* Tuning parameters (the autotuner's knobs; results never depend on them):
 *   "sort_items"   onesweep items per thread: 8, 12 (default) or 16 (tile = sort_block*items)
 *   "sort_block"   onesweep threads per workgroup: 256, 512 or 1024 (default;
 *                  1024 takes at most 12 items per thread)
 *   "digit_bits"   widest radix digit per onesweep pass, 4..8 (default 8)
 *   "hist_blocks"  workgroups of the digit-histogram pass, 1..4096 (default 4096)
 *   "merge_items"  merge-path steps per thread: 8, 12 or 16 (tile =
 *                  merge_block*items); 0 (default): 16
 *   "merge_block"  merge threads per workgroup: 256, 512 (default) or 1024
 *   "spgemm_small" longest output column (products) kept in LDS, 2..2048
 *   "narrow_keys"  1 (default): keys of at most 32 bits move as 32-bit words
 *                  between radix passes, and a merge tile whose keys span
 *                  less than 2^32 - 1 holds them in LDS as 32-bit offsets from
 *                  its smallest key; 0: always 64-bit
 *   "row_ranks"    1 (default): the lattice product's long output columns sort
 *                  the ranks of A's rows among the rows present instead of
 *                  the rows when that saves a radix pass or gives 32-bit keys
 *                  (one extra pass over A's keys); 0: rows
 *   "spgemm_pieces" 1 (default): long output columns are cut by row value into
 *                  pieces merged in LDS (the global sort takes only columns
 *                  with more than 64 B entries or rows clustered beyond a
 *                  piece); 0: every long column through the global sort
 *   "spgemm_single_copy" 1 (default): when output columns hold 64 or more
 *                  products on average, the columns of one B entry (a scaled
 *                  copy of one A column) are written by a kernel that deals
 *                  them to the XCDs by their place in A, so that each XCD's
 *                  L2 keeps the part of A it reads; 0: by the output kernel
 * No reference counterpart (the reference hard-codes radix 2048 and the
 * thresholds 3000/50, multi_index_array/LibMIARadix.h:454-457). */
int lnt_context_set_param(lnt_context ctx, const char *name, int64_t value);
int lnt_context_get_param(lnt_context ctx, const char *name, int64_t *value);
int lnt_context_set_profiling(lnt_context ctx, int on);
int lnt_context_profile(lnt_context ctx, const char *kernel, int64_t *launches,
                        double *total_ms, double *bytes);

/* ------------------------------------------------------- relinearisation -- */
/* In place: every key is recomputed from mode order cur_seq to new_seq.  Does
 * not sort.  Replaces SparseMIABase::change_linIdx_sequence
 * (multi_index_array/SparseMIABase.h:768-814, per-key
 * IndexUtil.h:567-589 reShuffleLinearIndex) and SparseChangeLinIdxMex
 * (MATLAB/mex/sparse_change_linIdx.cpp:17-71; that entry takes 1-based uint8
 * sequences, this one 0-based int32). */
int lnt_change_linidx(lnt_context ctx, int64_t *keys, int64_t n, int order,
                      const int64_t *dims, const int32_t *cur_seq,
                      const int32_t *new_seq, int mem);

/* ------------------------------------------------------------------ sort -- */
/* In place: sorts keys ascending and permutes vals alongside.  Keys must be
 * unique and in [0, max_size).  Replaces SparseMIABase::sort()
 * (multi_index_array/SparseMIABase.h:701-734 ->
 * LibMIARadix.h:444-484 RadixSortInPlace_PowerOf2Radix_Unsigned) and
 * SparseSortMex (MATLAB/mex/sparse_sort.cpp:14-68).  Equal keys keep their
 * input order.  From 2^32 entries on, the sort runs in chunks of 2^31: each
 * chunk ordered by the top 16 key bits, the chunks gathered by those bits,
 * then each group of about 2^31 entries sorted (16 bytes per entry of work
 * memory; LNT_ERR_TOO_LARGE if 2^32 keys share their top 16 bits). */
int lnt_sort_pairs(lnt_context ctx, int64_t *keys, double *vals, int64_t n,
                   uint64_t max_size, int mem);

/* ----------------------------------------------------- change_sort_order -- */
/* In place: relinearises every key from cur_seq to new_seq and sorts (keys,
 * vals) ascending in the new order.  `is_sorted` states whether the input is
 * sorted in cur_seq (a hint; the result does not depend on it).  Replaces
 * SparseMIABase::sort(seq) (multi_index_array/SparseMIABase.h:590-645, which
 * runs change_linIdx_sequence + RadixShuffle::permute, LibMIARadix.h:811-900)
 * and SparsePermuteMex (MATLAB/mex/sparse_permute.cpp:15-116, which expects the
 * keys already relinearised).  Any n (as lnt_sort_pairs). */
int lnt_change_sort_order(lnt_context ctx, int64_t *keys, double *vals,
                          int64_t n, int order, const int64_t *dims,
                          const int32_t *cur_seq, const int32_t *new_seq,
                          int is_sorted, int mem);

/* --------------------------------------------------- merge and Hadamard -- */
/* A and B sorted ascending (same mode order, unique keys).  op ADD/SUB: sorted
 * union, equal keys -> a op b, A only -> a, B only -> 0 op b (exact IEEE).  op
 * HADAMARD: sorted intersection with a*b.  C is sorted.  `capacity` must be at
 * least na+nb (ADD/SUB) or min(na,nb) (HADAMARD); *nc receives nnz(C).
 * Replaces outside_merge_sparse_storage_containers
 * (multi_index_array/FunctionUtil.h:373-419, reached from
 * SparseMIABase::outside_scanMerge, SparseMIABase.h:1404-1455), the pure
 * Hadamard case of perform_noLatticeMult (SparseMIABase.h:1717-1826) and
 * SparseMergeMex (MATLAB/mex/sparse_merge.cpp:15-162).  Inputs of 2^32
 * entries or more together are merged in chunks of at most 2^31 merge-path
 * steps, each cut so that a key present in both inputs stays in one chunk;
 * the cut is read back to the host once per chunk. */
int lnt_merge_into(lnt_context ctx, const int64_t *a_keys,
                   const double *a_vals, int64_t na, const int64_t *b_keys,
                   const double *b_vals, int64_t nb, int op, int64_t *c_keys,
                   double *c_vals, int64_t capacity, int64_t *nc, int mem);
/* lnt_merge_into for callers that cannot allocate (code recorded into a HIP
 * graph, tight serving loops, harnesses that forbid hipMalloc): no reference
 * counterpart.  Device memory only; the caller passes A, B and C as for
 * lnt_merge_into.  The call allocates nothing: its scratch (look-back words,
 * merge-path splits, a counter) is a static 512 KiB array that every GPU the
 * library runs on holds until the process ends, which limits na + nb to
 * 43,000 merge tiles (352,256,000 entries at the default 512 x 16 tile;
 * LNT_ERR_TOO_LARGE beyond).  All calls on one GPU share that array: calls
 * that overlap on one GPU (other threads, other streams) race. */
int lnt_merge_into_noalloc(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                           const int64_t *b_keys, const double *b_vals, int64_t nb, int op, int64_t *c_keys,
                           double *c_vals, int64_t capacity, int64_t *nc);
/* As lnt_merge_into, but the library allocates C (device memory for
 * LNT_MEM_DEVICE, host memory otherwise); release with lnt_free. */
int lnt_merge(lnt_context ctx, const int64_t *a_keys, const double *a_vals,
              int64_t na, const int64_t *b_keys, const double *b_vals,
              int64_t nb, int op, int64_t **c_keys, double **c_vals,
              int64_t *nc, int mem);
/* Union of two operands that are NOT sorted (SparseMIA::sortMerge,
 * SparseMIA.h:833-887, the path of C = A + B / A += B when neither operand is
 * sorted, SparseMIABase.h:1383-1398, SparseMIA.h:1030-1050).  Both key arrays
 * must already use one linearisation (the caller relinearises the smaller
 * operand, as sortMerge does) with keys in [0, bound).  A then B (B negated
 * for LNT_OP_SUB) are sorted together and equal keys collapsed by
 * collect_duplicates (SparseMIA.h:628-665): a key's first entry opens the key
 * only if its value is non-zero (the smallest key always opens), later
 * entries are added to it -- so keys whose entries are all zero vanish.  C is
 * sorted; capacity >= na + nb. */
int lnt_sort_merge_into(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                        const int64_t *b_keys, const double *b_vals, int64_t nb, int op, uint64_t bound,
                        int64_t *c_keys, double *c_vals, int64_t capacity, int64_t *nc, int mem);
/* As lnt_sort_merge_into with C allocated by the library (lnt_free). */
int lnt_sort_merge(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na,
                   const int64_t *b_keys, const double *b_vals, int64_t nb, int op, uint64_t bound,
                   int64_t **c_keys, double **c_vals, int64_t *nc, int mem);
/* Equality of two sparse tensors in one linearisation, both sorted ascending
 * with unique keys (SparseMIABase::compare_with_sparse,
 * SparseMIABase.h:1243-1296, after its sort of both operands into one order,
 * which the caller does).  Entries whose |value| <= 1.11e-16
 * (SparseTolerance<double>, LibMIAUtil.h:137-150; below_tolerance,
 * SparseMIABase.h:1170-1175) are skipped on either side; the other keys must
 * be the same on both sides and every pair must satisfy the predicate:
 * LNT_CMP_EXACT a == b (operator==, SparseMIABase.h:1616-1623), LNT_CMP_FUZZY
 * isEqualFuzzy(a, b, precision) (fuzzy_equals, SparseMIABase.h:1627-1635;
 * FunctionUtil.h:688-698).  *equal receives 1 or 0; nothing is written to the
 * operands.  Validation level 2 also checks that both are sorted. */
typedef enum lnt_compare_mode { LNT_CMP_EXACT = 0, LNT_CMP_FUZZY = 1 } lnt_compare_mode;
int lnt_compare(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na, const int64_t *b_keys,
                const double *b_vals, int64_t nb, int mode, double precision, int *equal, int mem);
int lnt_free(lnt_context ctx, void *ptr, int mem);
/* Copies `bytes` bytes from src to dst on the context's stream and waits.
 * Either side may be device memory (the library's or the caller's) or host
 * memory: how a caller moves a library-allocated result into its own
 * buffers before lnt_free, and how the C++ front end moves a tensor's
 * entries to and from HBM.  No counterpart in the reference (mex outputs are
 * mxArrays). */
int lnt_copy(lnt_context ctx, void *dst, const void *src, int64_t bytes);
/* *ptr receives `bytes` (>= 1) bytes of device memory (LNT_MEM_DEVICE) or
 * host memory (LNT_MEM_HOST) on the context's device, released with
 * lnt_free: storage for tensors kept resident in HBM between calls (the C++
 * front end's to_device).  No counterpart in the reference (its storage is
 * std::vector, SparseMIA.h:61-65). */
int lnt_alloc(lnt_context ctx, int64_t bytes, int mem, void **ptr);

/* ------------------------------------------------------- lattice product -- */
/* The accumulator of a lattice product, optionally or-ed with
 * LNT_SPGEMM_CSC.  Every strategy returns the same keys; SORT, DENSE, ESC and
 * MTT the same values bit for bit (each row summed in ascending B row order),
 * HASH values within the accumulation-order tolerance. */
typedef enum lnt_spgemm_strategy {
  LNT_SPGEMM_AUTO = 0,  /* library choice: SORT */
  LNT_SPGEMM_SORT = 1,  /* expand, sort rows, sum runs: csc_no_accum (SparseLatticeBase.h:1854-1990) */
  LNT_SPGEMM_HASH = 2,  /* LDS hash accumulation for short columns */
  LNT_SPGEMM_DENSE = 3, /* dense sparse accumulator over rows in LDS (csc_times, :2224-2351) when
                           h <= 8192, else SORT */
  LNT_SPGEMM_ESC = 4,   /* expand every product, one global sort, sum runs (outer_times, :2034-2183) */
  LNT_SPGEMM_MTT = 5,   /* dense accumulator over the ranks of A's rows present (mtt_times,
                           :1579-1813) when there are <= 8192 of them, else SORT */
  LNT_SPGEMM_CSC = 16   /* flag: find A's columns through a column-pointer array (CSC,
                           MultHelper<.,false>) instead of a search of A's keys (DCSC) */
} lnt_spgemm_strategy;

/* A lattice is an order-3 view of a tensor: dims {h, w, d} (rows, columns,
 * tabs), key = row + h*col + h*w*tab (multi_index_array/SparseLatticeBase.h:
 * 867-877).  A is h x w x d, B is w x bw x d, both sorted ascending by key.
 * *nprod receives the number of scalar products sum_{(k,j,t) in B}
 * nnz(A(:,k,t)) -- the bound on nnz(C) and half the flop count. */
int lnt_lattice_nprod(lnt_context ctx, const int64_t *a_keys, int64_t na,
                      const int64_t a_dims[3], const int64_t *b_keys,
                      int64_t nb, const int64_t b_dims[3], int64_t *nprod,
                      int mem);
/* C = A * B tab by tab (h x bw x d), sorted ascending; equal rows of a
 * column are summed, exact zeros kept.  Replaces SparseLatticeBase::operator*
 * -> csc_no_accum<true> (multi_index_array/SparseLatticeBase.h:1571-1576,
 * 1854-2031) and SparseLatticeMultMex (MATLAB/mex/sparselattice_mul.cpp:
 * 14-136).  Keys are bit-identical to the reference; values agree to 1e-12
 * (only the summation order of duplicate rows differs).  capacity >= nnz(C)
 * (lnt_lattice_nprod bounds it); at most 2^32 products per call. */
int lnt_lattice_mult_into(lnt_context ctx, const int64_t *a_keys,
                          const double *a_vals, int64_t na,
                          const int64_t a_dims[3], const int64_t *b_keys,
                          const double *b_vals, int64_t nb,
                          const int64_t b_dims[3], int strategy,
                          int64_t *c_keys, double *c_vals, int64_t capacity,
                          int64_t *nc, int mem);
/* As lnt_lattice_mult_into, with C allocated by the library at its exact size
 * (release with lnt_free). */
int lnt_lattice_mult(lnt_context ctx, const int64_t *a_keys,
                     const double *a_vals, int64_t na, const int64_t a_dims[3],
                     const int64_t *b_keys, const double *b_vals, int64_t nb,
                     const int64_t b_dims[3], int strategy, int64_t **c_keys,
                     double **c_vals, int64_t *nc, int mem);

/* ------------------------------------------ sparse x dense, no contraction -- */
/* C = A * B with B dense and no contracted index: C(l, r, t) = A(l, t) *
 * B(r, t) over A's outer indices l, B's outer indices r and the shared
 * element-wise indices t.  Replaces SparseMIABase::noLatticeMult with a
 * DenseMIA operand (multi_index_array/SparseMIABase.h:1829-1910).  A: keys
 * sorted ascending with key = l + l_size*t (its outer modes, then its inter
 * modes).  B: order_b modes of sizes b_dims, first mode fastest, host or device
 * memory like A; r_modes[nr] lists B's outer modes (r = their linear index,
 * first listed fastest), i_modes[ni] its inter modes in the order of A's
 * (t likewise); together they name each mode of B once.  For every A entry
 * with a non-zero value and every r whose B value is non-zero, in order, C
 * holds a*b at key r + l*R + t*l_size*R (R = prod of B's outer dims) -- sorted
 * ascending when A is.  C is library-allocated (lnt_free).  B holds at most
 * 2^40 elements. */
int lnt_sparse_dense_mult(lnt_context ctx, const int64_t *a_keys, const double *a_vals, int64_t na, int64_t l_size,
                          int order_b, const int64_t *b_dims, const double *b, const int32_t *r_modes, int nr,
                          const int32_t *i_modes, int ni, int64_t **c_keys, double **c_vals, int64_t *nc, int mem);

/* ------------------------------------- unary contraction / attraction -- */
typedef enum lnt_unary_kind {
  LNT_UNARY_CONTRACT = 0, /* sum over the diagonal of the group's modes */
  LNT_UNARY_ATTRACT = 1   /* keep the diagonal as one mode */
} lnt_unary_kind;

/* B = A with repeated indices resolved: A(i,j,i,k) (contraction) or
 * A(!i,j,!i) (attraction).  Replaces SparseMIABase::contract_attract
 * (multi_index_array/SparseMIABase.h:1913-2053), reached from an index
 * expression with a repeated index (MIA_Expr.h:695-731).  A is (keys, vals,
 * n) in mode order `seq`, any order of entries.  group[m] = 0: mode m stays;
 * group[m] = g in 1..ngroups: mode m belongs to group g, whose modes must have
 * one size; kind[g-1] says what group g does.  B's modes: the staying modes in
 * increasing mode order, then the attraction groups in increasing group
 * number; B's keys are in B's default order (first mode fastest), sorted
 * ascending, library-allocated (lnt_free).  Each B entry is the sum, from 0,
 * of the A entries on its diagonals in ascending order of their contraction
 * values, the highest-numbered contraction group most significant (the
 * reference's sort order when groups are numbered by the last occurrence of
 * their index, contractions before attractions); sums equal to zero are
 * dropped (:2044).  All modes contracted gives one entry at key 0 (or none). */
int lnt_contract_attract(lnt_context ctx, const int64_t *keys, const double *vals, int64_t n, int order,
                         const int64_t *dims, const int32_t *seq, const int32_t *group, const int32_t *kind,
                         int ngroups, int64_t **c_keys, double **c_vals, int64_t *nc, int mem);

#ifdef __cplusplus
}
#endif
#endif /* LNT_H */
