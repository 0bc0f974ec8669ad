/* lnt_dispatch.h -- one sparse contraction C = A * B on one MI355X (gfx950),
 * run by any of several strategies with identical results, and the learned
 * cost model that picks among them.
 *
 * lnt_contract is the whole expression C(...) = A(...) * B(...) of the
 * reference's front end: unfold both operands into lattices (MIA_Expr.h:
 * 339-419, 449-561; SparseMIABase.h:1457-1586), multiply (SparseLatticeBase.h:
 * 1574 operator*, csc_no_accum :1854-1990) and refold (SparseMIA.h:362-372;
 * MIA_Expr.h:1114-1165).  The reference chooses an orientation and an
 * algorithm per product from sparsity flags and the operands' current sort
 * orders (MATLAB sparse_times_polyalgorithm.m:1-208); here a strategy is
 * LNT_STRATEGY(accumulator, radix6, swap):
 *   swap        compute B*A and re-sort the result into C's order (the
 *               polyalgorithm's `transposed`);
 *   accumulator LNT_ACC_SORT  expand, sort rows, sum runs (csc_no_accum,
 *                             SparseLatticeBase.h:1854-1990);
 *               LNT_ACC_HASH  LDS hash accumulator for short columns;
 *               LNT_ACC_DENSE dense sparse accumulator (csc_times,
 *                             :2224-2351) when the lattice height is <= 8192;
 *               LNT_ACC_ESC   expand every product, one sort, sum runs
 *                             (outer_times, :2034-2183);
 *               LNT_ACC_MTT   dense accumulator over the ranks of A's rows
 *                             present (mtt_times, :1579-1813) when <= 8192;
 *               (DENSE and MTT use SORT where their accumulator would not
 *               fit in LDS);
 *   radix6      6-bit radix digits in the operand sorts instead of 8.
 * Every strategy returns the same keys bit for bit; SORT, DENSE, ESC and MTT
 * in one orientation the same values bit for bit, the others values within
 * the accumulation-order tolerance.  LNT_STRAT_AUTO asks the context's model
 * (lnt_context_set_model), or, without one, the restated polyalgorithm
 * (lnt_contract_heuristic).
 *
 * All key/value pointers are device pointers; operands are not modified.
 */
#ifndef LNT_DISPATCH_H
#define LNT_DISPATCH_H

#include "lnt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LNT_STRAT_SWAP 1
#define LNT_STRAT_RADIX6 2
#define LNT_ACC_SORT 0
#define LNT_ACC_HASH 1
#define LNT_ACC_DENSE 2
#define LNT_ACC_ESC 3
#define LNT_ACC_MTT 4
#define LNT_NUM_ACCUMULATORS 5
#define LNT_STRATEGY(acc, radix6, swap) (((acc) << 2) | ((radix6) ? LNT_STRAT_RADIX6 : 0) | ((swap) ? LNT_STRAT_SWAP : 0))
#define LNT_STRATEGY_ACC(s) ((s) >> 2)
#define LNT_NUM_STRATEGIES (4 * LNT_NUM_ACCUMULATORS)
#define LNT_STRAT_AUTO (-1)
#define LNT_NUM_FEATURES 18

/* One sparse operand: int64 linear indices in the order `seq` (first entry
 * fastest, SparseMIABase.h:101-153) and FP64 values; `labels` names each
 * mode (NUL-terminated, one character per mode; a label shared by A and B is
 * contracted, an upper-case shared label is an inter index -- the
 * reference's !i, Index.h:54 -- and an unshared label is an outer index);
 * `sorted` says the keys are ascending. */
typedef struct {
  int32_t order;
  int64_t dims[LNT_MAX_ORDER];
  int32_t seq[LNT_MAX_ORDER];
  char labels[LNT_MAX_ORDER + 1];
  int32_t sorted;
  int64_t nnz;
  const int64_t *keys;
  const double *vals;
} lnt_operand;

/* C's shape: modes [A-outer, B-outer, inter] (A's mode order within each
 * group), keys in the default sequence, ascending. */
typedef struct {
  int32_t order;
  int64_t dims[LNT_MAX_ORDER];
  char labels[LNT_MAX_ORDER + 1];
} lnt_result_shape;

/* C = A * B with `strategy` (0 .. LNT_NUM_STRATEGIES-1 or LNT_STRAT_AUTO).  The library allocates
 * c_keys/c_vals (free with lnt_free, LNT_MEM_DEVICE).  *used_strategy (may
 * be NULL) receives the strategy that ran.  Synchronises. */
int lnt_contract(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int strategy,
                 int64_t **c_keys, double **c_vals, int64_t *nnz_c, lnt_result_shape *shape,
                 int *used_strategy);

/* As lnt_contract, with C written to caller-allocated device arrays of
 * `capacity` entries (no allocation per call: for loops over many
 * contractions).  Fails with LNT_ERR_INVALID_ARGUMENT, writing nothing past
 * the arrays, when nnz(C) exceeds the capacity; nnz(C) <= the scalar product
 * count (lnt_lattice_nprod), and a pure Hadamard product needs min(nnz(A),
 * nnz(B)).  Synchronises. */
int lnt_contract_into(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int strategy,
                      int64_t *c_keys, double *c_vals, int64_t capacity, int64_t *nnz_c,
                      lnt_result_shape *shape, int *used_strategy);

/* The LNT_NUM_FEATURES features of the cost model (descriptors only, no data
 * is read): log2 of nnz(A), nnz(B), the lattice shape h, w, bw, d; the two
 * orders; contracted and inter mode counts; log2 of the estimated products
 * nnz(A)*nnz(B)/(w*d) and of the estimated nnz(C); whether A and B are
 * already sorted in the lattice order of each orientation (4 flags); log2 of
 * both dimensionalities. */
int lnt_contract_features(const lnt_operand *a, const lnt_operand *b, float *features);

/* The restated polyalgorithm (sparse_times_polyalgorithm.m): orientation from
 * the operands' current orders and hyper-sparsity flags (dim/nnz > 3), and its
 * algorithm class: 1 and 2 (csc_times CSC / DCSC) -> LNT_ACC_DENSE, 0 and 3
 * (csc_no_accum) -> LNT_ACC_SORT, 4 (outer product) -> LNT_ACC_ESC; 8-bit
 * digits. */
int lnt_contract_heuristic(const lnt_operand *a, const lnt_operand *b, int *strategy);

/* Wall time (ms, host clock around the call and a final synchronisation) of
 * each strategy in `mask` (bit s = strategy s), the minimum over `reps` runs
 * after one untimed warm-up run of that strategy.  The result goes to one
 * reusable buffer (sized by a first, untimed run), so no allocation is timed.
 * Unmeasured entries of ms[LNT_NUM_STRATEGIES] get -1. */
int lnt_contract_time(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int mask, int reps,
                      double *ms);

/* ------------------------------------------------------------- cost model
 * A 3-layer perceptron (LNT_NUM_FEATURES -> 64 -> 64 -> LNT_NUM_STRATEGIES,
 * ReLU, FP32) predicting log2(ms) of every strategy from standardised
 * features; trained on the GPU by the library's own forward/backward
 * kernels with Adam on the mean squared error over measured entries. */
typedef struct lnt_model_s *lnt_model;

#define LNT_MODEL_HIDDEN 64

int lnt_model_create(lnt_context ctx, uint64_t seed, lnt_model *out);
int lnt_model_destroy(lnt_model model);
/* Number of float parameters (weights, biases) plus 2*LNT_NUM_FEATURES
 * standardisation constants (mean, 1/std) stored after them. */
int64_t lnt_model_param_count(void);
int lnt_model_get_params(lnt_model model, float *params);
int lnt_model_set_params(lnt_model model, const float *params);
/* Host arrays: features n x LNT_NUM_FEATURES, targets n x LNT_NUM_STRATEGIES
 * (log2 ms; NaN where unmeasured).  Sets the standardisation from `features`,
 * then trains.  *loss (may be NULL) receives the final full-set loss. */
int lnt_model_train(lnt_model model, const float *features, const float *targets, int64_t n, int epochs,
                    int batch, float learning_rate, uint64_t seed, float *loss);
/* out: n x LNT_NUM_STRATEGIES predicted log2 ms. */
int lnt_model_predict(lnt_model model, const float *features, int64_t n, float *out);
/* Loss and its gradient with respect to every weight and bias (the first
 * lnt_model_param_count() - 2*LNT_NUM_FEATURES entries of `grad`) over the
 * given set with the current parameters, by the training kernels. */
int lnt_model_gradient(lnt_model model, const float *features, const float *targets, int64_t n, float *grad,
                       float *loss);
/* The model used by LNT_STRAT_AUTO in this context (NULL: the heuristic).
 * The context does not own it. */
int lnt_context_set_model(lnt_context ctx, lnt_model model);

#ifdef __cplusplus
}
#endif

#endif /* LNT_DISPATCH_H */
