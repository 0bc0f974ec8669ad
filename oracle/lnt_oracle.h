/*
 * lnt_oracle.h -- CPU restatement of LibMIA's sparse multi-index-array hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline; it is never linked into, loaded by, or called from the
 * product library (libnt_amd/liblnt.so).  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.
 *
 * Conventions (all follow the reference, paths relative to src/multi_index_array/):
 *   - a sparse tensor is two parallel arrays: int64 linear indices ("keys",
 *     always >= 0) and double values (SparseMIA.h:61-65, 203-206);
 *   - a key is computed in a mode order "seq" whose first entry is the fastest
 *     varying mode: key = sum_t sub[seq[t]] * prod_{u<t} dims[seq[u]]
 *     (IndexUtil.h:866-879 sub2ind_reorder, IndexUtil.h:318-333 ind2sub_reorder);
 *   - "sorted" means ascending keys; keys of one tensor are unique.
 *
 * Parity status of each function is recorded in the header comment of its
 * definition in lnt_oracle.c and pinned by tests/test_oracle.py.
 */
#ifndef LNT_ORACLE_H
#define LNT_ORACLE_H

#define LNTO_MAX_ORDER 16

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- bit-portable synthetic-input generators ---------------- */
uint64_t lnto_mix64(uint64_t x);
uint64_t lnto_rand_u64(uint64_t seed, uint64_t i);
double lnto_uniform_pm1(uint64_t seed, uint64_t i);
double lnto_uniform_01(uint64_t seed, uint64_t i);
double lnto_normal(uint64_t seed, uint64_t i);
uint64_t lnto_feistel(uint64_t x, int bits, uint64_t seed);
uint64_t lnto_permute_index(uint64_t i, uint64_t domain, uint64_t seed);
void lnto_gen_distinct_keys(int64_t *keys, int64_t n, uint64_t domain,
                            uint64_t seed, uint64_t offset);
void lnto_gen_uniform_vals(double *vals, int64_t n, uint64_t seed,
                           uint64_t offset);
void lnto_random_perm(int32_t *perm, int n, uint64_t seed);
/* Draws distinct coordinates mode by mode (dist[m]: 0 uniform, 1 Zipf with
 * exponent zipf_s over [0,dims[m])), rejecting duplicates, values N(0,1)
 * (value_kind 1) or U[-1,1] (value_kind 0).  Keys in the default seq. */
int lnto_gen_coords(int64_t *keys, double *vals, int64_t nnz, int order,
                    const int64_t *dims, const int32_t *dist, double zipf_s,
                    int value_kind, uint64_t seed);

/* ------------------------------ index math ------------------------------ */
int64_t lnto_sub2ind(const int64_t *sub, const int32_t *seq,
                     const int64_t *dims, int order);
void lnto_ind2sub(int64_t key, const int32_t *seq, const int64_t *dims,
                  int order, int64_t *sub);
void lnto_change_linidx(int64_t *keys, int64_t n, int order,
                        const int64_t *dims, const int32_t *cur_seq,
                        const int32_t *new_seq);

/* -------------------------------- sorting ------------------------------- */
void lnto_radix_sort_pairs(int64_t *keys, double *vals, int64_t n,
                           uint64_t max_size);
void lnto_introsort_pairs(int64_t *keys, double *vals, int64_t n);
/* sort(seq) by the reference's own path for sorted input of 3000+ keys:
 * relinearise, then RadixShuffle::permute (LibMIARadix.h:811-900).  Equal to
 * lnto_sort_seq for power-of-two dims (the CPU baseline's path). */
void lnto_sort_seq_shuffle(int64_t *keys, double *vals, int64_t n, int order, const int64_t *dims,
                           const int32_t *cur_seq, const int32_t *new_seq, int is_sorted);
void lnto_sort_seq(int64_t *keys, double *vals, int64_t n, int order,
                   const int64_t *dims, const int32_t *cur_seq,
                   const int32_t *new_seq, int is_sorted);
int64_t lnto_collect_duplicates(int64_t *keys, double *vals, int64_t n);
/* SparseMIA::collect_duplicates(Collector = plus) (SparseMIA.h:628-665) on
 * sorted keys: a key's first entry is kept only if its value is non-zero
 * (the array's first entry always), later entries of the kept key are added
 * to it.  Returns the new length. */
int64_t lnto_collect_duplicates_member(int64_t *keys, double *vals, int64_t n);
/* SparseMIA::sortMerge (SparseMIA.h:833-887) on operands already in one
 * linearisation: concatenate A then B (B negated, -1*b, for LNTO_OP_SUB),
 * sort (the reference sort of sort()), collect_duplicates_member.  ck/cv hold
 * na+nb entries; returns nnz(C). */
int64_t lnto_sort_merge(const int64_t *ak, const double *av, int64_t na, const int64_t *bk,
                        const double *bv, int64_t nb, int op, uint64_t bound, int64_t *ck,
                        double *cv);

/* -------------------------------- merging ------------------------------- */
enum { LNTO_OP_ADD = 0, LNTO_OP_SUB = 1, LNTO_OP_HADAMARD = 2 };
int64_t lnto_merge(const int64_t *ak, const double *av, int64_t na,
                   const int64_t *bk, const double *bv, int64_t nb, int op,
                   int64_t *ck, double *cv);
/* SparseMergeMex: lnto_merge (op 0/1) behind the mex's empty-operand cases. */
int64_t lnto_mex_merge(const int64_t *ak, const double *av, int64_t na,
                       const int64_t *bk, const double *bv, int64_t nb, int op,
                       int64_t *ck, double *cv);
int64_t lnto_outer_inter_count(const int64_t *ak, int64_t na,
                               int64_t l_outer_size, const int64_t *bk,
                               int64_t nb, int64_t r_outer_size);
int64_t lnto_outer_inter_product(const int64_t *ak, const double *av,
                                 int64_t na, int64_t l_outer_size,
                                 const int64_t *bk, const double *bv,
                                 int64_t nb, int64_t r_outer_size,
                                 int64_t *ck, double *cv);

/* --------------------------- lattice product ---------------------------- */
int64_t lnto_lattice_nprod(const int64_t *ak, int64_t na, int64_t ah,
                           int64_t aw, const int64_t *bk, int64_t nb,
                           int64_t bh, int64_t bw);
int64_t lnto_lattice_mult(const int64_t *ak, const double *av, int64_t na,
                          int64_t ah, int64_t aw, int64_t ad,
                          const int64_t *bk, const double *bv, int64_t nb,
                          int64_t bh, int64_t bw, int64_t bd, int64_t *ck,
                          double *cv, int64_t cap);
/* csc_times (compress = 0) / mtt_times (compress = 1) sparse-accumulator
 * products and the outer_times expand-sort-collect product of the same
 * lattices (both sorted column major); nnz(C) or -1. */
int64_t lnto_lattice_mult_spa(const int64_t *ak, const double *av, int64_t na, int64_t ah, int64_t aw,
                              const int64_t *bk, const double *bv, int64_t nb, int64_t bh, int64_t bw,
                              int compress, int64_t *ck, double *cv, int64_t cap);
int64_t lnto_lattice_mult_outer(const int64_t *ak, const double *av, int64_t na, int64_t ah, int64_t aw,
                                const int64_t *bk, const double *bv, int64_t nb, int64_t bh, int64_t bw,
                                int64_t *ck, double *cv, int64_t cap);
/* Full MIA contraction A(labels_a) * B(labels_b) (lattice product path of
 * MIA_Expr.h:864-925).  labels are ASCII mode labels; an upper-case label is an
 * inter ("!") index.  Sorts A and B in place to their lattice orders.  Writes C
 * in C's default seq over dims [A-outer, B-outer, inter]; returns nnz(C) or -1. */
int64_t lnto_contract(int64_t *ak, double *av, int64_t na, int order_a,
                      const int64_t *dims_a, const int32_t *seq_a,
                      int a_sorted, const char *labels_a, int64_t *bk,
                      double *bv, int64_t nb, int order_b,
                      const int64_t *dims_b, const int32_t *seq_b,
                      int b_sorted, const char *labels_b, int64_t *ck,
                      double *cv, int64_t cap, int64_t *dims_c, int *order_c);
int lnto_contract_plan(int order_a, const int64_t *dims_a, const char *labels_a,
                       int order_b, const int64_t *dims_b, const char *labels_b,
                       int32_t *seq_a, int32_t *seq_b, int64_t *lat_a,
                       int64_t *lat_b, int64_t *dims_c, int *order_c);

/* ------------------------------ cost model ------------------------------ *
 * FP64 restatement of the dispatch model (no reference counterpart): a
 * perceptron F=18 -> 64 -> 64 -> S=20 with ReLU (derivative 0 at 0); params
 * W1[64][18] b1[64] W2[64][64] b2[64] W3[20][64] b3[20] then mean[18] and
 * inv_std[18]; input x is standardised as (x - mean) * inv_std.  Loss =
 * (1/m) sum over measured targets (not NaN) of (y - t)^2, m = their count. */
#define LNTO_MLP_F 18
#define LNTO_MLP_H 64
#define LNTO_MLP_S 20
int64_t lnto_mlp_param_count(void);   /* trainable + standardisation */
int64_t lnto_mlp_weight_count(void);  /* trainable only */
void lnto_mlp_forward(const double *params, const double *x, double *y);
/* Returns the loss; grad receives d loss / d (trainable params). */
double lnto_mlp_loss_grad(const double *params, const double *X, const double *T, int64_t n,
                          double *grad);

/* -------------------------- sparse x dense ----------------------------- */
/* SparseMIABase.h:1829-1910 noLatticeMult with a dense B (see lnt_oracle.c).
 * ck/cv need room for na * prod(B outer dims) entries. */
int64_t lnto_sparse_dense_mult(const int64_t *ak, const double *av, int64_t na, int64_t L, int order_b,
                               const int64_t *b_dims, const double *b, const int32_t *r_modes, int nr,
                               const int32_t *i_modes, int ni, int64_t *ck, double *cv);

/* ------------------------ unary contraction ---------------------------- */
/* SparseMIABase.h:1913-2053 contract_attract (see lnt_oracle.c).  ck/cv need
 * room for n entries; returns the number written, sorted by key.  Sorts
 * keys/vals in place. */
int64_t lnto_contract_attract(int64_t *keys, double *vals, int64_t n, int order, const int64_t *dims,
                              const int32_t *seq, const int32_t *group, const int32_t *kind, int ngroups,
                              int64_t *ck, double *cv);

/* ------------------------------- comparison ----------------------------- */
int lnto_is_equal_fuzzy(double a, double b, double prec);
/* compare_with_sparse (SparseMIABase.h:1243-1296) on two operands already
 * sorted in one linearisation: mode 0 exact (operator==, :1616-1623), mode 1
 * isEqualFuzzy with prec (fuzzy_equals, :1627-1635).  Returns 1 or 0. */
int lnto_compare_sparse(const int64_t *ak, const double *av, int64_t na, const int64_t *bk, const double *bv,
                        int64_t nb, int mode, double prec);

#ifdef __cplusplus
}
#endif
#endif
