// contract.hip -- one contraction C = A * B by a chosen strategy, the
// features of the cost model, the restated polyalgorithm, and strategy
// timing (include/lnt_dispatch.h).
//
// A strategy is (orientation, accumulator, radix width).  The unfold follows
// MIA_Expr.h:449-561: the first operand's lattice has rows = its outer modes,
// columns = contracted modes, tabs = inter modes, each in its own mode order;
// the second's rows = contracted modes matched to the first's order, columns =
// its outer modes, tabs = inter modes.  The swapped orientation computes
// B * A and re-sorts the result into A * B's mode order
// ([A-outer, B-outer, inter], the refold of SparseMIA.h:362-372), the
// `transposed` case of sparse_times_polyalgorithm.m.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>

#include "../../include/lnt_dispatch.h"
#include "lnt_internal.h"

namespace lnt {
int model_predict_one_host(lnt_model m, const float *features, float *y);
}

namespace lnt {

namespace {

struct Plan {
  int n_out_a = 0, n_in = 0, n_int = 0, n_out_b = 0;
  int a_out[LNT_MAX_ORDER], a_in[LNT_MAX_ORDER], a_int[LNT_MAX_ORDER];
  int b_in[LNT_MAX_ORDER], b_out[LNT_MAX_ORDER], b_int[LNT_MAX_ORDER];
  int32_t seq_a[LNT_MAX_ORDER], seq_b[LNT_MAX_ORDER];
  int64_t lat_a[3], lat_b[3];
  lnt_result_shape shape;
};

bool is_upper(char c) { return c >= 'A' && c <= 'Z'; }

int check_operand(const lnt_operand *x, const char *name) {
  const std::string n(name);
  if (!x) return fail(LNT_ERR_INVALID_ARGUMENT, "null operand " + n);
  if (x->order < 1 || x->order > LNT_MAX_ORDER) return fail(LNT_ERR_INVALID_ARGUMENT, n + ": order out of range");
  if ((int)strnlen(x->labels, LNT_MAX_ORDER + 1) != x->order)
    return fail(LNT_ERR_INVALID_ARGUMENT, n + ": one label per mode is required");
  if (x->nnz < 0 || (x->nnz > 0 && (!x->keys || !x->vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, n + ": bad nnz or null arrays");
  if (x->nnz >= (int64_t(1) << 32)) return fail(LNT_ERR_TOO_LARGE, n + ": operands hold < 2^32 entries");
  for (int m = 0; m < x->order; ++m) {
    if (x->dims[m] < 1) return fail(LNT_ERR_INVALID_ARGUMENT, n + ": dims must be >= 1");
    for (int r = m + 1; r < x->order; ++r)
      if (x->labels[m] == x->labels[r]) return fail(LNT_ERR_INVALID_ARGUMENT, n + ": a label repeats");
  }
  return LNT_OK;
}

int find_label(const lnt_operand *x, char c) {
  for (int m = 0; m < x->order; ++m)
    if (x->labels[m] == c) return m;
  return -1;
}

int64_t prod_of(const lnt_operand *x, const int *modes, int n, unsigned __int128 *acc) {
  unsigned __int128 p = 1;
  for (int i = 0; i < n; ++i) p *= (uint64_t)x->dims[modes[i]];
  *acc = p;
  return (int64_t)p;
}

int make_plan(const lnt_operand *a, const lnt_operand *b, Plan *p) {
  Plan P;
  for (int m = 0; m < a->order; ++m) {
    const char c = a->labels[m];
    const int r = find_label(b, c);
    if (r >= 0) {
      if (a->dims[m] != b->dims[r]) return fail(LNT_ERR_DIMENSION_MISMATCH, std::string("index ") + c + " differs in size");
      if (is_upper(c)) {
        P.a_int[P.n_int] = m;
        P.b_int[P.n_int++] = r;
      } else {
        P.a_in[P.n_in] = m;
        P.b_in[P.n_in++] = r;
      }
    } else {
      if (is_upper(c)) return fail(LNT_ERR_INVALID_ARGUMENT, std::string("inter index ") + c + " in one operand only");
      P.a_out[P.n_out_a++] = m;
    }
  }
  for (int r = 0; r < b->order; ++r) {
    const char c = b->labels[r];
    if (find_label(a, c) >= 0) continue;
    if (is_upper(c)) return fail(LNT_ERR_INVALID_ARGUMENT, std::string("inter index ") + c + " in one operand only");
    P.b_out[P.n_out_b++] = r;
  }
  if (P.n_out_a + P.n_out_b + P.n_int == 0)
    return fail(LNT_ERR_INVALID_ARGUMENT, "a full contraction to a scalar is not a sparse tensor");
  int q = 0;
  for (int i = 0; i < P.n_out_a; ++i) P.seq_a[q++] = P.a_out[i];
  for (int i = 0; i < P.n_in; ++i) P.seq_a[q++] = P.a_in[i];
  for (int i = 0; i < P.n_int; ++i) P.seq_a[q++] = P.a_int[i];
  q = 0;
  for (int i = 0; i < P.n_in; ++i) P.seq_b[q++] = P.b_in[i];
  for (int i = 0; i < P.n_out_b; ++i) P.seq_b[q++] = P.b_out[i];
  for (int i = 0; i < P.n_int; ++i) P.seq_b[q++] = P.b_int[i];
  unsigned __int128 h, w, d, bw;
  P.lat_a[0] = prod_of(a, P.a_out, P.n_out_a, &h);
  P.lat_a[1] = prod_of(a, P.a_in, P.n_in, &w);
  P.lat_a[2] = prod_of(a, P.a_int, P.n_int, &d);
  P.lat_b[0] = P.lat_a[1];
  P.lat_b[1] = prod_of(b, P.b_out, P.n_out_b, &bw);
  P.lat_b[2] = P.lat_a[2];
  if (h * bw * d > ((unsigned __int128)1 << 63)) return fail(LNT_ERR_TOO_LARGE, "C's dimensionality exceeds 2^63");
  lnt_result_shape &S = P.shape;
  memset(&S, 0, sizeof(S));
  for (int i = 0; i < P.n_out_a; ++i) {
    S.dims[S.order] = a->dims[P.a_out[i]];
    S.labels[S.order++] = a->labels[P.a_out[i]];
  }
  for (int i = 0; i < P.n_out_b; ++i) {
    S.dims[S.order] = b->dims[P.b_out[i]];
    S.labels[S.order++] = b->labels[P.b_out[i]];
  }
  for (int i = 0; i < P.n_int; ++i) {
    S.dims[S.order] = a->dims[P.a_int[i]];
    S.labels[S.order++] = a->labels[P.a_int[i]];
  }
  *p = P;
  return LNT_OK;
}

bool same_seq(const lnt_operand *x, const int32_t *seq) {
  for (int m = 0; m < x->order; ++m)
    if (x->seq[m] != seq[m]) return false;
  return true;
}

// The operand in the lattice sequence `lat_seq` (change_sort_order,
// SparseMIABase.h:590-645): a sorted operand already in that sequence is used
// where it lies; otherwise it is copied to the context's scratch and sorted
// there.  *keys / *vals receive the arrays to read.
int unfold(lnt_context ctx, const lnt_operand *x, const int32_t *lat_seq, DeviceBuffer &kb, DeviceBuffer &vb,
           const uint64_t **keys, const double **vals) {
  Relin r;
  uint64_t dim;
  LNT_TRY(make_relin(x->order, x->dims, x->seq, lat_seq, &r, &dim));
  if (x->nnz == 0 || (x->sorted && !r.active)) {
    *keys = (const uint64_t *)x->keys;
    *vals = x->vals;
    return LNT_OK;
  }
  const size_t bytes = (size_t)x->nnz * 8;
  LNT_TRY(ensure(kb, bytes));
  LNT_TRY(ensure(vb, bytes));
  LNT_HIP(hipMemcpyAsync(kb.ptr, x->keys, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  LNT_HIP(hipMemcpyAsync(vb.ptr, x->vals, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  *keys = (const uint64_t *)kb.ptr;
  *vals = (const double *)vb.ptr;
  const uint32_t low = x->sorted ? sorted_low_bits(x->order, x->dims, x->seq, lat_seq) : 0;
  return sort_device(ctx, (uint64_t *)kb.ptr, (double *)vb.ptr, (uint32_t)x->nnz, dim, &r, low);
}

struct DigitScope {
  lnt_context ctx;
  int saved;
  DigitScope(lnt_context c, int bits) : ctx(c), saved(c->tuning.digit_bits) { ctx->tuning.digit_bits = bits; }
  ~DigitScope() { ctx->tuning.digit_bits = saved; }
};

// Runs strategy s.  With out_k/out_v (capacity entries) the result is
// written there; otherwise it is allocated with hipMalloc.
int run_strategy(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int s, uint64_t *out_k,
                 double *out_v, uint64_t capacity, uint64_t **ck, double **cv, uint64_t *nc,
                 lnt_result_shape *shape, uint64_t *nprod = nullptr) {
  *ck = nullptr;
  *cv = nullptr;
  *nc = 0;
  const bool fixed = out_k != nullptr;
  const bool swap = s & LNT_STRAT_SWAP;
  const lnt_operand *f = swap ? b : a, *g = swap ? a : b;
  Plan P, P0;
  LNT_TRY(make_plan(f, g, &P));
  LNT_TRY(make_plan(a, b, &P0));
  *shape = P0.shape;
  DigitScope digits(ctx, (s & LNT_STRAT_RADIX6) ? 6 : 8);
  static const int kSpgemm[LNT_NUM_ACCUMULATORS] = {LNT_SPGEMM_SORT, LNT_SPGEMM_HASH, LNT_SPGEMM_DENSE,
                                                     LNT_SPGEMM_ESC, LNT_SPGEMM_MTT};
  const uint64_t *fk = nullptr, *gk = nullptr;
  const double *fv = nullptr, *gv = nullptr;
  LNT_TRY(unfold(ctx, f, P.seq_a, ctx->exec_buf[0], ctx->exec_buf[1], &fk, &fv));
  LNT_TRY(unfold(ctx, g, P.seq_b, ctx->exec_buf[2], ctx->exec_buf[3], &gk, &gv));
  uint64_t *k = out_k;
  double *v = out_v;
  uint64_t n = 0;
  auto release_out = [&]() {
    if (!fixed) {
      if (k) (void)hipFree(k);
      if (v) (void)hipFree(v);
    }
  };
  if (P.n_out_a == 0 && P.n_in == 0 && P.n_out_b == 0) {
    // Pure Hadamard (SparseMIABase.h:1717-1826): both operands now share one
    // linearisation; sorted intersection with products.
    const uint64_t cap = (uint64_t)std::min(f->nnz, g->nnz);
    if (fixed) {
      if (capacity < cap) return fail(LNT_ERR_INVALID_ARGUMENT, "output capacity below min(nnz(A), nnz(B))");
    } else if (hipMalloc((void **)&k, (cap ? cap : 1) * 8) != hipSuccess ||
               hipMalloc((void **)&v, (cap ? cap : 1) * 8) != hipSuccess) {
      (void)hipGetLastError();
      if (k) (void)hipFree(k);
      return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc of the Hadamard product failed");
    }
    uint32_t c = 0;
    int rc = cap ? merge_device(ctx, fk, fv, (uint32_t)f->nnz, gk, gv, (uint32_t)g->nnz, LNT_OP_HADAMARD, k, v, cap, &c)
                 : LNT_OK;
    if (rc != LNT_OK) {
      release_out();
      return rc;
    }
    n = c;
  } else {
    const int acc = kSpgemm[LNT_STRATEGY_ACC(s)];
    int rc = fixed ? lattice_mult_device(ctx, fk, fv, (uint64_t)f->nnz, P.lat_a, gk, gv, (uint64_t)g->nnz, P.lat_b,
                                         acc, k, v, capacity, &n, nprod, nullptr, nullptr)
                   : lattice_mult_device(ctx, fk, fv, (uint64_t)f->nnz, P.lat_a, gk, gv, (uint64_t)g->nnz, P.lat_b,
                                         acc, nullptr, nullptr, 0, &n, nprod, &k, &v);
    if (rc != LNT_OK) return rc;
    if (!fixed && !k) {  // empty product: nothing allocated
      if (hipMalloc((void **)&k, 8) != hipSuccess || hipMalloc((void **)&v, 8) != hipSuccess) {
        (void)hipGetLastError();
        if (k) (void)hipFree(k);
        return fail(LNT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
      }
    }
  }
  if (swap && n > 0 && P.shape.order > 1) {
    // C' has modes [B-outer, A-outer, inter (B's order)]: relinearise into
    // C's mode order and sort.
    int32_t cur[LNT_MAX_ORDER], nxt[LNT_MAX_ORDER];
    for (int m = 0; m < P.shape.order; ++m) {
      cur[m] = m;
      nxt[m] = -1;
      for (int r = 0; r < P.shape.order; ++r)
        if (P.shape.labels[r] == P0.shape.labels[m]) nxt[m] = r;
    }
    Relin r;
    uint64_t dim;
    int rc = make_relin(P.shape.order, P.shape.dims, cur, nxt, &r, &dim);
    // C' comes out sorted in its own order.
    if (rc == LNT_OK && r.active)
      rc = sort_device(ctx, k, v, (uint32_t)n, dim, &r, sorted_low_bits(P.shape.order, P.shape.dims, cur, nxt));
    if (rc != LNT_OK) {
      release_out();
      return rc;
    }
  }
  *ck = k;
  *cv = v;
  *nc = n;
  return LNT_OK;
}

float lg(double x) { return (float)log2(x > 0 ? x : 1.0); }

int features_of(const lnt_operand *a, const lnt_operand *b, float *f) {
  Plan ab, ba;
  LNT_TRY(make_plan(a, b, &ab));
  LNT_TRY(make_plan(b, a, &ba));
  const double h = (double)ab.lat_a[0], w = (double)ab.lat_a[1], d = (double)ab.lat_a[2], bw = (double)ab.lat_b[1];
  const double na = (double)a->nnz, nb = (double)b->nnz;
  const double nprod = na * nb / (w * d);
  const double cells = h * bw * d;
  const double nnz_c = cells * -expm1(-nprod / cells);
  double dim_a = 1, dim_b = 1;
  for (int m = 0; m < a->order; ++m) dim_a *= (double)a->dims[m];
  for (int m = 0; m < b->order; ++m) dim_b *= (double)b->dims[m];
  f[0] = lg(na + 1);
  f[1] = lg(nb + 1);
  f[2] = lg(h);
  f[3] = lg(w);
  f[4] = lg(bw);
  f[5] = lg(d);
  f[6] = (float)a->order;
  f[7] = (float)b->order;
  f[8] = (float)ab.n_in;
  f[9] = (float)ab.n_int;
  f[10] = lg(nprod + 1);
  f[11] = lg(nnz_c + 1);
  f[12] = (a->sorted && same_seq(a, ab.seq_a)) ? 1.f : 0.f;
  f[13] = (b->sorted && same_seq(b, ab.seq_b)) ? 1.f : 0.f;
  f[14] = (a->sorted && same_seq(a, ba.seq_b)) ? 1.f : 0.f;
  f[15] = (b->sorted && same_seq(b, ba.seq_a)) ? 1.f : 0.f;
  f[16] = lg(dim_a);
  f[17] = lg(dim_b);
  return LNT_OK;
}

// sparse_times_polyalgorithm.m restated on the lattice plan: lexorder tests
// compare the leading (inner + outer) entries of each operand's current
// sequence with [outer inner] / [inner outer]; hyper-sparsity flags are
// size/nnz > 3; `transposed` -> LNT_STRAT_SWAP.  Its algorithm classes map to
// the accumulators of the methods they name (specialized_mult.m and
// sparselattice_mul.cpp:24-67): 1 CSC and 2 DCSC (csc_times) -> DENSE, 0
// operator* and 3 CSCNA (csc_no_accum) -> SORT, 4 outer product -> ESC
// (orientation A*B, as the .m file's `transposed = 2` lattices are); 8-bit
// digits.
int heuristic_of(const lnt_operand *a, const lnt_operand *b, int *strategy) {
  Plan P;
  LNT_TRY(make_plan(a, b, &P));
  const double na = a->nnz > 0 ? (double)a->nnz : 1.0, nb = b->nnz > 0 ? (double)b->nnz : 1.0;
  const double a_inner = (double)P.lat_a[1], a_outer = (double)P.lat_a[0];
  const double b_inner = (double)P.lat_b[0], b_outer = (double)P.lat_b[1];
  bool a_col = a_inner / na > 3, a_row = a_outer / na > 3;
  const bool a_sparse = !(a_col || a_row), a_hyper = a_col && a_row;
  a_row = a_row && !a_hyper;
  a_col = a_col && !a_hyper;
  bool b_row = b_inner / nb > 3, b_col = b_outer / nb > 3;
  const bool b_sparse = !(b_col || b_row), b_hyper = b_col && b_row;
  b_row = b_row && !b_hyper;
  b_col = b_col && !b_hyper;
  // lexorder tests on the leading entries of the current sequences
  auto lead_is = [](const lnt_operand *x, const int *g1, int n1, const int *g2, int n2) {
    for (int i = 0; i < n1; ++i)
      if (x->seq[i] != g1[i]) return false;
    for (int i = 0; i < n2; ++i)
      if (x->seq[n1 + i] != g2[i]) return false;
    return true;
  };
  const bool a_oi = lead_is(a, P.a_out, P.n_out_a, P.a_in, P.n_in);
  const bool a_io = lead_is(a, P.a_in, P.n_in, P.a_out, P.n_out_a);
  const bool b_io = lead_is(b, P.b_in, P.n_in, P.b_out, P.n_out_b);
  const bool b_oi = lead_is(b, P.b_out, P.n_out_b, P.b_in, P.n_in);
  // The six-way lexorder choice shared by the CSC, DCSC and CSCNA branches.
  // *fifth: the fifth case was taken (the .m file's sparse x sparse branch
  // runs algorithm 0 there, :50-52).
  auto orient = [&](int *transposed, bool *fifth) {
    *fifth = false;
    if (a_oi && b_io) *transposed = 0;
    else if (a_io && b_oi) *transposed = 1;
    else if (a_oi && !b_oi) *transposed = 0;
    else if (!a_oi && b_oi) *transposed = 1;
    else if (a_io && !b_io) *transposed = 1, *fifth = true;
    else if (!a_io && b_io) *transposed = 0;
    else *transposed = b_outer > a_outer ? 0 : 1;
  };
  int transposed = 0, algorithm = 1;
  bool fifth = false;
  if (a_sparse) {
    algorithm = 1;
    if (b_sparse) {
      orient(&transposed, &fifth);
      if (fifth) algorithm = 0;
    } else {
      transposed = 0;
    }
  } else if (b_sparse) {
    transposed = 1;
    algorithm = 1;
  } else if (a_col) {
    if (b_row) orient(&transposed, &fifth);
    else transposed = 0;
    algorithm = 2;
  } else if (b_row) {
    transposed = 1;
    algorithm = 2;
  } else if (a_row) {
    if (b_col) orient(&transposed, &fifth);
    else transposed = 0;
    algorithm = 3;
  } else if (b_col) {
    transposed = 1;
    algorithm = 3;
  } else {
    transposed = 2;
    algorithm = 4;
  }
  static const int kAcc[5] = {LNT_ACC_SORT, LNT_ACC_DENSE, LNT_ACC_DENSE, LNT_ACC_SORT, LNT_ACC_ESC};
  *strategy = LNT_STRATEGY(kAcc[algorithm], 0, transposed == 1);
  return LNT_OK;
}

}  // namespace
}  // namespace lnt

using namespace lnt;

extern "C" {

int lnt_contract_features(const lnt_operand *a, const lnt_operand *b, float *features) {
  if (!features) return fail(LNT_ERR_INVALID_ARGUMENT, "null features");
  LNT_TRY(check_operand(a, "A"));
  LNT_TRY(check_operand(b, "B"));
  return features_of(a, b, features);
}

int lnt_contract_heuristic(const lnt_operand *a, const lnt_operand *b, int *strategy) {
  if (!strategy) return fail(LNT_ERR_INVALID_ARGUMENT, "null strategy");
  LNT_TRY(check_operand(a, "A"));
  LNT_TRY(check_operand(b, "B"));
  return heuristic_of(a, b, strategy);
}

static int choose(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int *strategy) {
  if (!ctx->model) return heuristic_of(a, b, strategy);
  float f[LNT_NUM_FEATURES], y[LNT_NUM_STRATEGIES];
  LNT_TRY(features_of(a, b, f));
  LNT_TRY(model_predict_one_host((lnt_model)ctx->model, f, y));
  int best = 0;
  for (int s = 1; s < LNT_NUM_STRATEGIES; ++s)
    if (y[s] < y[best]) best = s;
  *strategy = best;
  return LNT_OK;
}

int lnt_contract(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int strategy, int64_t **c_keys,
                 double **c_vals, int64_t *nnz_c, lnt_result_shape *shape, int *used_strategy) {
  if (!ctx || !c_keys || !c_vals || !nnz_c || !shape) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  *c_keys = nullptr;
  *c_vals = nullptr;
  LNT_TRY(check_operand(a, "A"));
  LNT_TRY(check_operand(b, "B"));
  if (strategy != LNT_STRAT_AUTO && (strategy < 0 || strategy >= LNT_NUM_STRATEGIES))
    return fail(LNT_ERR_INVALID_ARGUMENT, "unknown strategy");
  LNT_HIP(hipSetDevice(ctx->device));
  if (strategy == LNT_STRAT_AUTO) LNT_TRY(choose(ctx, a, b, &strategy));
  uint64_t *k;
  double *v;
  uint64_t n;
  LNT_TRY(run_strategy(ctx, a, b, strategy, nullptr, nullptr, 0, &k, &v, &n, shape));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  *c_keys = (int64_t *)k;
  *c_vals = v;
  *nnz_c = (int64_t)n;
  if (used_strategy) *used_strategy = strategy;
  return LNT_OK;
}

int lnt_contract_into(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int strategy, int64_t *c_keys,
                      double *c_vals, int64_t capacity, int64_t *nnz_c, lnt_result_shape *shape, int *used_strategy) {
  if (!ctx || !nnz_c || !shape) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  if (capacity < 0 || (capacity > 0 && (!c_keys || !c_vals)))
    return fail(LNT_ERR_INVALID_ARGUMENT, "negative capacity or null output arrays");
  *nnz_c = 0;
  LNT_TRY(check_operand(a, "A"));
  LNT_TRY(check_operand(b, "B"));
  if (strategy != LNT_STRAT_AUTO && (strategy < 0 || strategy >= LNT_NUM_STRATEGIES))
    return fail(LNT_ERR_INVALID_ARGUMENT, "unknown strategy");
  LNT_CTX(ctx);
  if (strategy == LNT_STRAT_AUTO) LNT_TRY(choose(ctx, a, b, &strategy));
  uint64_t *k;
  double *v;
  uint64_t n;
  // A one-element dummy keeps run_strategy's caller-buffer mode for an empty
  // capacity: any non-empty result then fails the capacity check.
  uint64_t dummy_k;
  double dummy_v;
  uint64_t *ok = capacity > 0 ? (uint64_t *)c_keys : &dummy_k;
  double *ov = capacity > 0 ? c_vals : &dummy_v;
  LNT_TRY(run_strategy(ctx, a, b, strategy, ok, ov, (uint64_t)capacity, &k, &v, &n, shape));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  *nnz_c = (int64_t)n;
  if (used_strategy) *used_strategy = strategy;
  return LNT_OK;
}

int lnt_contract_time(lnt_context ctx, const lnt_operand *a, const lnt_operand *b, int mask, int reps,
                      double *ms) {
  if (!ctx || !ms || reps < 1) return fail(LNT_ERR_INVALID_ARGUMENT, "bad arguments");
  LNT_TRY(check_operand(a, "A"));
  LNT_TRY(check_operand(b, "B"));
  LNT_HIP(hipSetDevice(ctx->device));
  for (int s = 0; s < LNT_NUM_STRATEGIES; ++s) ms[s] = -1.0;
  if (!(mask & ((1 << LNT_NUM_STRATEGIES) - 1))) return LNT_OK;
  // One allocating run gives nnz(C) and the number of scalar products (the
  // same for every strategy); every timed run then writes into one reusable
  // buffer of that many entries, so no allocation falls inside a timed
  // region and the lattice product writes its long columns straight into it.
  uint64_t *k;
  double *v;
  uint64_t n, nprod = 0;
  lnt_result_shape shape;
  LNT_TRY(run_strategy(ctx, a, b, 0, nullptr, nullptr, 0, &k, &v, &n, &shape, &nprod));
  LNT_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(k);
  (void)hipFree(v);
  const uint64_t cap = std::max<uint64_t>({n, nprod, (uint64_t)std::min(a->nnz, b->nnz), 1});
  LNT_TRY(ensure(ctx->exec_buf[4], cap * 8));
  LNT_TRY(ensure(ctx->exec_buf[5], cap * 8));
  uint64_t *ok = (uint64_t *)ctx->exec_buf[4].ptr;
  double *ov = (double *)ctx->exec_buf[5].ptr;
  for (int s = 0; s < LNT_NUM_STRATEGIES; ++s) {
    if (!((mask >> s) & 1)) continue;
    double best = 1e300;
    for (int rep = 0; rep <= reps; ++rep) {  // rep 0 warms this strategy's workspaces up
      LNT_HIP(hipStreamSynchronize(ctx->stream));
      const auto t0 = std::chrono::steady_clock::now();
      LNT_TRY(run_strategy(ctx, a, b, s, ok, ov, cap, &k, &v, &n, &shape));
      LNT_HIP(hipStreamSynchronize(ctx->stream));
      const auto t1 = std::chrono::steady_clock::now();
      const double t = std::chrono::duration<double, std::milli>(t1 - t0).count();
      if (rep > 0 && t < best) best = t;
    }
    ms[s] = best;
  }
  return LNT_OK;
}

}  // extern "C"
