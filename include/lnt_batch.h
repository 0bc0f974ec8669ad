/* lnt_batch.h -- many independent sparse contractions in one pass on one
 * MI355X (gfx950), with the whole pipeline recordable as a hipGraph.
 *
 * The reference evaluates one expression at a time: every product
 * C(...) = A(...) * B(...) unfolds both operands into lattices
 * (MIA_Expr.h:339-419, SparseMIABase.h:1457-1586 toStraightLattice /
 * toLatticeSort), multiplies them (SparseLatticeBase.h:1574 operator*,
 * csc_no_accum :1854-1990) and refolds the result (SparseMIA.h:362-372).
 * A batch does exactly that for every instance, with all instances sharing
 * each kernel launch; it has no counterpart call in the reference.  Each
 * instance's result equals lnt_lattice_mult (lnt.h) on the same operands after
 * lnt_change_sort_order into lattice order: keys bit-exact, values within the
 * accumulation-order tolerance.
 *
 * Memory: all pointers are device pointers.  Operands are packed: instance i's
 * A keys/values start at sum_{j<i} nnz_a[j], likewise B.  Keys are int64
 * linear indices in the instance's current sequence (cur_seq), as in
 * SparseMIA (SparseMIABase.h:101-153).
 */
#ifndef LNT_BATCH_H
#define LNT_BATCH_H

#include "lnt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lnt_batch_s *lnt_batch;

/* One contraction C = A * B in lattice form.  new_seq_a orders A's modes as
 * [A-outer, contracted, inter] and new_seq_b orders B's as [contracted (in
 * A's order), B-outer, inter] -- the unfold of MIA_Expr.h:449-561.  lat_a =
 * (h, w, d) and lat_b = (w, bw, d) are the lattice shapes (rows, columns, tabs);
 * C is h x bw x d with key = row + h*col + h*bw*tab, i.e. C's modes are
 * [A-outer, B-outer, inter] in the default sequence. */
typedef struct {
  int32_t order_a, order_b;
  int64_t dims_a[LNT_MAX_ORDER], dims_b[LNT_MAX_ORDER];
  int32_t cur_seq_a[LNT_MAX_ORDER], new_seq_a[LNT_MAX_ORDER];
  int32_t cur_seq_b[LNT_MAX_ORDER], new_seq_b[LNT_MAX_ORDER];
  int64_t lat_a[3], lat_b[3];
  int64_t nnz_a, nnz_b;
} lnt_contraction;

/* Validates every instance and uploads the batch description.  The batch
 * keeps `ctx` (and its stream) for all later calls. */
int lnt_batch_create(lnt_context ctx, const lnt_contraction *items, int64_t count, lnt_batch *out);
int lnt_batch_destroy(lnt_batch batch);

/* Runs every contraction.  The library allocates the packed result (free both
 * arrays with lnt_free); c_offsets (host, count+1 entries) receives the start
 * of each instance's C inside it, c_offsets[count] = total nnz.  Synchronises. */
int lnt_batch_contract(lnt_batch batch, const int64_t *a_keys, const double *a_vals,
                       const int64_t *b_keys, const double *b_vals, int64_t **c_keys,
                       double **c_vals, int64_t *c_offsets);

/* The same in two phases, for callers that allocate the result themselves:
 * the symbolic phase (sorts, column tables, per-column output counts) returns
 * nnz(C) of the whole batch; the numeric phase then writes C into c_keys /
 * c_vals (capacity >= nnz_c entries) and c_offsets (host, count+1, may be
 * NULL).  The numeric phase needs the symbolic phase of the same batch
 * immediately before it (no other call on this batch in between). */
int lnt_batch_symbolic(lnt_batch batch, const int64_t *a_keys, const double *a_vals,
                       const int64_t *b_keys, const double *b_vals, int64_t *nnz_c);
int lnt_batch_numeric(lnt_batch batch, int64_t *c_keys, double *c_vals, int64_t capacity,
                      int64_t *c_offsets);

/* Records the whole pipeline for these buffers into a hipGraph (one graph
 * per batch; a new capture replaces the old one).  The result is written to
 * c_keys/c_vals (capacity entries).  The batch's workspaces are sized by one
 * unrecorded run first, which also returns the exact result size in *nnz_c
 * (may be NULL); a replay on operands with more output than `capacity`
 * reports LNT_ERR_TOO_LARGE and writes nothing past it.  The graph's sort
 * scratch (ping-pong buffers, look-back words, histograms) belongs to the
 * batch, not to the context: later calls on the context, lnt_context_trim
 * included, leave a recorded graph valid; the batch frees it on destroy. */
int lnt_batch_capture(lnt_batch batch, const int64_t *a_keys, const double *a_vals,
                      const int64_t *b_keys, const double *b_vals, int64_t *c_keys,
                      double *c_vals, int64_t capacity, int64_t *nnz_c);

/* Launches the recorded graph and waits for it.  c_offsets (host, count+1,
 * may be NULL) as for lnt_batch_contract.  The operand buffers given to
 * lnt_batch_capture may hold new contents of the same sizes. */
int lnt_batch_replay(lnt_batch batch, int64_t *c_offsets);

/* Launches the recorded graph without waiting (for timing loops); the next
 * lnt_batch_replay or lnt_context_synchronize completes it. */
int lnt_batch_launch(lnt_batch batch);

/* Tuning of the batch pipeline (results never depend on it):
 *   "item_products"  products per small-segment work item (64..1024)
 *   "small_max"      largest segment grouped into items   (2..1024)
 *   "window"         target products per window of a large segment (64..1536)
 *   "blocks_per_cu"  persistent workgroups per compute unit (1..8)
 *   "piece_slots"    at most this many row pieces per run (-1: no limit, the
 *                    default; 0: every large segment in windows).  A large
 *                    segment of at most 64 B entries is cut by row value into
 *                    pieces of about 3,072 products merged in LDS; the pool of
 *                    piece slots grows to what the last run needed, and a
 *                    segment that finds no room is processed in windows.
 *   "dense_min"      smallest segment (products) of a dense instance that
 *                    takes the dense accumulator (16..2^30; default 256)
 * The sort parameters are the context's (lnt_context_set_param). */
int lnt_batch_set_param(lnt_batch batch, const char *name, int64_t value);
/* lnt_batch_get_param also reads "dense_segments": the output columns the
 * dense accumulator took in the last run (see lnt_batch_set_accumulators). */
int lnt_batch_get_param(lnt_batch batch, const char *name, int64_t *value);

/* The accumulator of each instance (count entries, LNT_ACC_* of
 * lnt_dispatch.h), the per-contraction choice of the strategy set inside one
 * batch.  LNT_ACC_DENSE and LNT_ACC_MTT: the instance's output columns of two
 * or more B entries and at least "dense_min" products (8 or more per entry)
 * are summed in a dense accumulator over the lattice rows in LDS -- the
 * sparse accumulator of csc_times (SparseLatticeBase.h:2224-2351) -- when its
 * lattice height is at most 8192 (otherwise the choice is ignored); every
 * other value: the batch's merge and sort paths.  Results are bit-identical
 * whatever the choice (every row summed in ascending B-row order), except a
 * row met by more than 2,048 products of one output column, which the sort
 * paths sum by a tree (within the accumulation-order tolerance); only the
 * time differs.  Applies to later calls and to replays of a recorded graph.
 * Default after lnt_batch_create: LNT_ACC_SORT everywhere.  Synchronises. */
int lnt_batch_set_accumulators(lnt_batch batch, const int32_t *acc);

/* Exact scalar-product count of every instance (sum over k of
 * nnzA(:,k) * nnzB(k,:)); products: host array of count entries.  Runs the
 * sort and column phases only.  Synchronises. */
int lnt_batch_nprod(lnt_batch batch, const int64_t *a_keys, const int64_t *b_keys, int64_t *products);

#ifdef __cplusplus
}
#endif

#endif /* LNT_BATCH_H */
