/*
 * lnt_mex.h -- the reference's MATLAB mex entry points, restated as plain C
 * functions over the gfx950 library.
 *
 * The reference's FFI for its sparse hot path is five mex files
 * (src/MATLAB/mex/sparse_sort.cpp, sparse_change_linIdx.cpp,
 * sparse_permute.cpp, sparse_merge.cpp, sparselattice_mul.cpp).  Each
 * function below takes the arrays a mexFunction receives (mxGetData of its
 * right-hand sides, host memory) and gives them the same meaning: 1-based
 * uint8 mode sequences, in-place mutation of the caller's arrays where the
 * mex mutates its inputs, and freshly allocated outputs where the mex returns
 * new mxArrays.  A mexFunction around one of them checks the mxArray classes
 * and shapes (the reference's check_sparse.cpp:7-275), calls it, and turns a
 * non-zero status into mexErrMsgTxt(lnt_last_error()).
 *
 * Conventions
 *   - Data is double (mxDOUBLE_CLASS, the NTToolbox's class).  The sorts only
 *     move values, never compute with them, so any 8-byte class passed as
 *     double* comes back bit for bit.
 *   - Indices are int64 linear indices, 0-based (the MATLAB side subtracts 1
 *     before the call, @SparseNT/changeLexOrder.m:10-13), >= 0.
 *   - Mode sequences are uint8 and 1-based, as MATLAB passes them; dims are
 *     int64, one per mode.
 *   - Outputs that the mex returns as new arrays are allocated with `alloc`,
 *     called as alloc(count, element_size) -- pass mxCalloc from a mex file,
 *     or NULL for calloc (release with free) -- after the result is known,
 *     and only for a non-empty result: an empty result hands back NULL
 *     pointers and a count of 0 (the reference's 0x0 outputs).  Nothing is
 *     allocated or written on error.
 *   - Every function returns an lnt_status (lnt.h); lnt_last_error() gives
 *     the reason.  All return with the work finished (the context's stream is
 *     synchronised).
 */
#ifndef LNT_MEX_H
#define LNT_MEX_H

#include <stddef.h>
#include <stdint.h>

#include "lnt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mxCalloc's shape: count elements of `size` bytes, zeroed. */
typedef void *(*lnt_mex_alloc)(size_t count, size_t size);

/* SparseSortMex(data, idx) (MATLAB/mex/sparse_sort.cpp:15-20, 23-68): sorts
 * idx ascending in place and permutes data alongside.  Unique indices have
 * one sorted order, which the radix sort gives.  The reference sorts with
 * Introsort (unstable), so equal indices end in the order its swaps leave:
 * when the input holds equal indices and n <= 65,536, the GPU replays that
 * Introsort step for step (bit-identical data order); above that, equal
 * indices keep their input order.  n < 2^32. */
int lnt_mex_sparse_sort(lnt_context ctx, double *data, int64_t *idx, int64_t n);

/* SparseChangeLinIdxMex(idx, curSeq, newSeq, dims)
 * (MATLAB/mex/sparse_change_linIdx.cpp:17-71): recomputes every index in
 * place from mode order cur_seq to new_seq (1-based), without sorting.
 * Returns at once when n == 0 or the sequences are equal (:30-31, :49-50),
 * as the mex does; otherwise the sequences must be permutations of 1..order. */
int lnt_mex_sparse_change_linidx(lnt_context ctx, int64_t *idx, int64_t n, const uint8_t *cur_seq,
                                 const uint8_t *new_seq, const int64_t *dims, int order);

/* SparsePermuteMex(data, idx, curSeq, newSeq, dims)
 * (MATLAB/mex/sparse_permute.cpp:15-34, 39-116): idx already holds the
 * indices in the new_seq linearisation, in the order of a tensor sorted in
 * cur_seq; sorts (idx, data) ascending in place.  Returns at once when n == 0
 * or the sequences are equal (:57-58, :76-77).  The fastest modes of new_seq that keep their
 * relative order from cur_seq are already ordered, and the radix passes skip
 * their bits, as the reference's RadixShuffle skips its first stage
 * (multi_index_array/LibMIARadix.h:733-740).  n < 2^32. */
int lnt_mex_sparse_permute(lnt_context ctx, double *data, int64_t *idx, int64_t n, const uint8_t *cur_seq,
                           const uint8_t *new_seq, const int64_t *dims, int order);

/* SparseMergeMex(a_data, a_idx, b_data, b_idx, op)
 * (MATLAB/mex/sparse_merge.cpp:16-85, 116-162, check_sparse.cpp:105-127): both
 * operands sorted ascending; op is floored and must be 0 (add) or 1
 * (subtract).  C is the sorted union (lnt_merge_into).  As the mex: an empty
 * B returns a copy of A; an empty A returns B, or -1*b for subtraction
 * (sparse_merge.cpp:26-57, so +0.0 becomes -0.0 there, where the general
 * merge computes 0 - b).  *c_data / *c_idx receive arrays of *nc elements. */
int lnt_mex_sparse_merge(lnt_context ctx, const double *a_data, const int64_t *a_idx, int64_t na,
                         const double *b_data, const int64_t *b_idx, int64_t nb, double op, lnt_mex_alloc alloc,
                         double **c_data, int64_t **c_idx, int64_t *nc);

/* SparseLatticeMultMex(a_data, a_idx, a_dims, b_data, b_idx, b_dims[,
 * algorithm]) (MATLAB/mex/sparselattice_mul.cpp:15-92, 94-136): C = A*B tab
 * by tab, both lattices (dims {h, w, d}) sorted ascending; C sorted, exact
 * zeros kept (lnt_lattice_mult_into).  `algorithm` is truncated to an
 * integer as the mex does; 0..5 name the reference's methods, each run by
 * its GPU counterpart (lnt.h lnt_spgemm_strategy): 0 operator* =
 * csc_no_accum<true> -> SORT; 1 csc_times<false> (CSC) -> DENSE | CSC; 2
 * csc_times<true> (DCSC) -> DENSE; 3 csc_no_accum<false> (CSC) -> SORT | CSC;
 * 5 mtt_times -> MTT; 4 -> MTT as well, because the mex's case 4 falls
 * through into case 5 (sparselattice_mul.cpp:52-58), whose mtt_times result
 * on the same arguments is what it returns.  All give the same keys and the
 * same values bit for bit (each row summed in ascending B row order).  Other
 * values fail as the mex does.  *c_data / *c_idx receive arrays of *nc elements. */
int lnt_mex_sparselattice_mult(lnt_context ctx, const double *a_data, const int64_t *a_idx, int64_t na,
                               const int64_t a_dims[3], const double *b_data, const int64_t *b_idx, int64_t nb,
                               const int64_t b_dims[3], double algorithm, lnt_mex_alloc alloc, double **c_data,
                               int64_t **c_idx, int64_t *nc);

#ifdef __cplusplus
}
#endif
#endif /* LNT_MEX_H */
