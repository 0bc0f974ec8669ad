// SparseMIA.hpp -- header-only C++17 front end over the C-ABI of lnt.h.
//
// Mirrors the reference's sparse multi-index-array API so that code written
// against it keeps compiling:
//   NTINDEX i; NTINDEX j; ...                         (Index.h:27-30)
//   lnt::SparseMIA<double,4> A(32,32,32,32);          (SparseMIA.h:167-773)
//   C(i,j,m,n) = A(i,j,k,l) * B(k,l,m,n);              (MIA_Expr.h:864-925)
//   C(i,j,k,l) = A(i,j,k,l) + B(k,l,i,j);              (MIA_Expr.h:1185-1250)
//   C(i,j) = A(!i,!j) * B(!i,!j);                      Hadamard (MIA_Expr.h:1012-1059)
//   A.sort(seq); A.change_linIdx_sequence(seq);        (SparseMIABase.h:590-814)
// Storage is the reference's: a std::vector of int64 linear indices and one of
// values, computed in the mode order linIdxSequence().  Every operation runs
// on the GPU through lnt.h (host arrays are staged to HBM and back); only
// double values are supported.  Errors become exceptions, as in the reference
// (LibMIAException.h:25-92).
//
// A tensor can instead keep its entries in HBM (A.to_device(), or
// A.assign_device() with library-allocated arrays): every operation on it
// then runs on the device arrays and its results stay in HBM too, so a chain
// of expressions moves no data over PCIe; A.to_host() brings the entries
// back.  An operation that meets one operand in HBM and one on the host moves
// the host one to HBM first.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../lnt.h"

namespace lnt {

// ------------------------------------------------------------- exceptions --
struct MIAException : std::runtime_error {
  int status;
  MIAException(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};
struct DimensionMismatchException : MIAException {
  using MIAException::MIAException;
};
struct MIAParameterException : MIAException {
  using MIAException::MIAException;
};

inline void check(int status) {
  if (status == LNT_OK) return;
  std::string msg = lnt_last_error();
  if (status == LNT_ERR_DIMENSION_MISMATCH) throw DimensionMismatchException(status, msg);
  if (status == LNT_ERR_INVALID_ARGUMENT) throw MIAParameterException(status, msg);
  throw MIAException(status, msg);
}

// ---------------------------------------------------------------- context --
class Context {
 public:
  explicit Context(int device = 0, void *stream = nullptr) { check(lnt_context_create(&ctx_, device, stream)); }
  ~Context() { lnt_context_destroy(ctx_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  lnt_context get() const { return ctx_; }

 private:
  lnt_context ctx_ = nullptr;
};

// One context per host thread (the C-ABI contexts are not shared).
inline Context &default_context() {
  thread_local std::unique_ptr<Context> ctx;
  if (!ctx) ctx.reset(new Context(0));
  return *ctx;
}

// ------------------------------------------------------------------ index --
// An index label.  `!i` marks an inter-product (element-wise) index; each `!`
// adds one to its element-wise count (Index.h:38-57), which `~` lowers again
// after a product (MIA_Expr.h:257-301, 1376).
template <size_t ID, int ElemWise = 0>
struct ProdInd {
  static constexpr size_t id = ID;
  static constexpr int elemval = ElemWise;
  ProdInd<ID, ElemWise + 1> operator!() const { return {}; }
};
struct Label {
  size_t id;
  int bangs;  // element-wise count: 0 = plain index
  bool inter() const { return bangs > 0; }
  bool operator==(const Label &o) const { return id == o.id && bangs == o.bangs; }
};

template <class T, size_t N>
class SparseMIA;

namespace detail {

inline int64_t product(const std::vector<int64_t> &v) {
  int64_t p = 1;
  for (int64_t x : v) p *= x;
  return p;
}

// Device memory the library allocated (lnt_alloc, or a result it handed
// over), released with lnt_free when the last owner goes.  The free needs no
// context, so tensors that outlive their thread's context stay safe.
struct DeviceArray {
  void *p = nullptr;
  explicit DeviceArray(void *q) : p(q) {}
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() {
    if (p) lnt_free(nullptr, p, LNT_MEM_DEVICE);
  }
  static std::shared_ptr<DeviceArray> make(int64_t bytes) {
    void *q = nullptr;
    check(lnt_alloc(default_context().get(), bytes > 0 ? bytes : 1, LNT_MEM_DEVICE, &q));
    return std::make_shared<DeviceArray>(q);
  }
  // A device copy of `bytes` bytes at src (host or device memory).
  static std::shared_ptr<DeviceArray> copy_of(const void *src, int64_t bytes) {
    auto d = make(bytes);
    if (bytes > 0) check(lnt_copy(default_context().get(), d->p, src, bytes));
    return d;
  }
};

// Type-erased tensor used by expressions (any order).  Its entries live in
// idx/data (host) or, when `device`, in dk/dv (dn entries in HBM).
struct Tensor {
  std::vector<int64_t> dims;
  std::vector<int32_t> seq;
  std::vector<int64_t> idx;
  std::vector<double> data;
  bool sorted = false;
  bool device = false;
  std::shared_ptr<DeviceArray> dk, dv;
  int64_t dn = 0;

  Tensor() = default;
  Tensor(Tensor &&) = default;
  Tensor &operator=(Tensor &&) = default;
  // A copy owns copies of the entries, in the same memory.
  Tensor(const Tensor &o)
      : dims(o.dims), seq(o.seq), idx(o.idx), data(o.data), sorted(o.sorted), device(o.device), dn(o.dn) {
    if (device) {
      dk = DeviceArray::copy_of(o.dk->p, dn * 8);
      dv = DeviceArray::copy_of(o.dv->p, dn * 8);
    }
  }
  Tensor &operator=(const Tensor &o) {
    if (this != &o) *this = Tensor(o);
    return *this;
  }

  int mem() const { return device ? LNT_MEM_DEVICE : LNT_MEM_HOST; }
  int64_t nnz() const { return device ? dn : (int64_t)idx.size(); }
  int64_t *keys() { return device ? (int64_t *)dk->p : idx.data(); }
  double *vals() { return device ? (double *)dv->p : data.data(); }

  // Takes over the entries of a library-allocated result in memory kind m.
  void adopt(int64_t *ck, double *cv, int64_t nc, int m) {
    if (m == LNT_MEM_DEVICE) {
      std::vector<int64_t>().swap(idx);
      std::vector<double>().swap(data);
      dk = std::make_shared<DeviceArray>(ck);
      dv = std::make_shared<DeviceArray>(cv);
      dn = nc;
      device = true;
      return;
    }
    idx.assign(ck, ck + nc);
    data.assign(cv, cv + nc);
    lnt_free(default_context().get(), ck, LNT_MEM_HOST);
    lnt_free(default_context().get(), cv, LNT_MEM_HOST);
    dk.reset();
    dv.reset();
    dn = 0;
    device = false;
  }
  // Moves the entries to HBM (no-op when already there).
  void to_device() {
    if (device) return;
    const int64_t n = (int64_t)idx.size();
    dk = DeviceArray::copy_of(idx.data(), n * 8);
    dv = DeviceArray::copy_of(data.data(), n * 8);
    dn = n;
    device = true;
    std::vector<int64_t>().swap(idx);
    std::vector<double>().swap(data);
  }
  // Moves the entries back to host vectors.
  void to_host() {
    if (!device) return;
    idx.resize((size_t)dn);
    data.resize((size_t)dn);
    if (dn) {
      check(lnt_copy(default_context().get(), idx.data(), dk->p, dn * 8));
      check(lnt_copy(default_context().get(), data.data(), dv->p, dn * 8));
    }
    dk.reset();
    dv.reset();
    dn = 0;
    device = false;
  }

  // change_linIdx_sequence (SparseMIABase.h:768-814): relinearise, unsorted.
  void relinearise(const std::vector<int32_t> &new_seq) {
    if (new_seq == seq) return;
    if (nnz())
      check(lnt_change_linidx(default_context().get(), keys(), nnz(), (int)dims.size(), dims.data(), seq.data(),
                              new_seq.data(), mem()));
    seq = new_seq;
    sorted = false;
  }

  void sort_to(const std::vector<int32_t> &new_seq) {
    if (nnz())
      check(lnt_change_sort_order(default_context().get(), keys(), vals(), nnz(), (int)dims.size(), dims.data(),
                                  seq.data(), new_seq.data(), sorted ? 1 : 0, mem()));
    seq = new_seq;
    sorted = true;
  }
};

// Operands of one call in one memory: a host operand meeting one in HBM moves
// there.
inline void same_memory(Tensor &a, Tensor &b) {
  if (a.device != b.device) (a.device ? b : a).to_device();
}

struct Expr {
  std::shared_ptr<Tensor> t;  // owning result or a view of a SparseMIA's storage
  std::vector<Label> labels;
  // Inter indices a product left at the back of `labels` (the reference's
  // inter_product_number, MIA_Expr.h:903-920): the ones `~` decrements.
  size_t inter_count = 0;
};

inline std::vector<int32_t> identity(size_t n) {
  std::vector<int32_t> s(n);
  for (size_t i = 0; i < n; ++i) s[i] = (int32_t)i;
  return s;
}

inline int find_label(const std::vector<Label> &ls, size_t id) {
  for (size_t i = 0; i < ls.size(); ++i)
    if (ls[i].id == id) return (int)i;
  return -1;
}

// outside_merge (SparseMIABase.h:1357-1455): outside_scanMerge when an operand
// is sorted, sortMerge when neither is.
inline Expr merge(Tensor &a, const std::vector<Label> &la, Tensor &b_in, const std::vector<Label> &lb, int op) {
  if (la.size() != lb.size()) throw DimensionMismatchException(LNT_ERR_DIMENSION_MISMATCH, "merge orders differ");
  Tensor b_copy;
  Tensor *bp = &b_in;
  if (&a == &b_in) {
    b_copy = b_in;
    bp = &b_copy;
  }
  Tensor &b = *bp;
  same_memory(a, b);
  const size_t n = la.size();
  std::vector<int32_t> index_order(n), lhs_order(n);
  for (size_t m = 0; m < n; ++m) {
    int r = find_label(lb, la[m].id);
    if (r < 0 || la[m].inter() || lb[r].inter())
      throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "merge operands need the same plain indices");
    if (a.dims[m] != b.dims[r]) throw DimensionMismatchException(LNT_ERR_DIMENSION_MISMATCH, "merge dims differ");
    index_order[m] = r;
    lhs_order[r] = (int32_t)m;
  }
  std::vector<int32_t> b_in_a(n), a_in_b(n);
  for (size_t t = 0; t < n; ++t) b_in_a[t] = index_order[a.seq[t]];
  for (size_t t = 0; t < n; ++t) a_in_b[t] = lhs_order[b.seq[t]];
  if (a.sorted && b.sorted) {
    if (b.nnz() < a.nnz()) b.sort_to(b_in_a);
    else a.sort_to(a_in_b);
  } else if (b.sorted) {
    a.sort_to(a_in_b);
  } else if (a.sorted) {
    b.sort_to(b_in_a);
  } else {
    // Neither sorted: C = copy of A, C.sortMerge(B) (SparseMIABase.h:1393-1398,
    // SparseMIA.h:833-887): the smaller operand is relinearised into the
    // other's sequence, then concatenation, sort and collect_duplicates.
    auto c = std::make_shared<Tensor>();
    c->dims = a.dims;
    c->sorted = true;
    // A's keys relinearised into a copy (A itself is left as it is).
    std::vector<int64_t> a_idx;
    std::shared_ptr<DeviceArray> a_dk;
    int64_t *ak = a.keys();
    if (a.nnz() > b.nnz()) {
      b.relinearise(b_in_a);
      c->seq = a.seq;
    } else {
      if (a.nnz() && a.seq != a_in_b) {
        if (a.device) {
          a_dk = DeviceArray::copy_of(a.keys(), a.nnz() * 8);
          ak = (int64_t *)a_dk->p;
        } else {
          a_idx = a.idx;
          ak = a_idx.data();
        }
        check(lnt_change_linidx(default_context().get(), ak, a.nnz(), (int)n, a.dims.data(), a.seq.data(),
                                a_in_b.data(), a.mem()));
      }
      c->seq = a_in_b;
    }
    int64_t *ck = nullptr;
    double *cv = nullptr;
    int64_t nc = 0;
    check(lnt_sort_merge(default_context().get(), ak, a.vals(), a.nnz(), b.keys(), b.vals(), b.nnz(), op,
                         (uint64_t)product(a.dims), &ck, &cv, &nc, a.mem()));
    c->adopt(ck, cv, nc, a.mem());
    return {c, la};
  }
  auto c = std::make_shared<Tensor>();
  c->dims = a.dims;
  c->seq = a.seq;
  c->sorted = true;
  int64_t *ck = nullptr;
  double *cv = nullptr;
  int64_t nc = 0;
  check(lnt_merge(default_context().get(), a.keys(), a.vals(), a.nnz(), b.keys(), b.vals(), b.nnz(), op, &ck, &cv,
                  &nc, a.mem()));
  c->adopt(ck, cv, nc, a.mem());
  return {c, la};
}

// Unary contraction / attraction of an operand indexed with a repeated index,
// A(i,j,i,k) or A(!i,j,!i) (perform_unary, MIA_Expr.h:695-731 ->
// SparseMIABase::contract_attract, SparseMIABase.h:1913-2053).  A repeated
// plain index is contracted, a repeated `!` index attracted.  Groups are
// numbered as the reference's expression code pulls them: contractions, then
// attractions, each in the order of the last occurrence of its index
// (ExprUtil.h:120-240) -- which fixes the order of the sums.  The result's
// labels: the other indices in mode order, then the attracted ones.
inline Expr unary(const std::shared_ptr<Tensor> &ap, const std::vector<Label> &la) {
  Tensor &a = *ap;
  const size_t n = la.size();
  std::vector<size_t> con, att;  // positions of the last occurrences
  for (size_t m = 0; m < n; ++m) {
    size_t count = 0, last = m;
    for (size_t q = 0; q < n; ++q)
      if (la[q].id == la[m].id) {
        ++count;
        last = q;
        if (la[q].bangs != la[m].bangs)
          throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "a repeated index has different element-wise counts");
      }
    if (count > 1 && last == m) (la[m].inter() ? att : con).push_back(m);
  }
  if (con.empty() && att.empty()) return Expr{ap, la};
  std::vector<int32_t> group(n, 0), kind;
  std::vector<Label> lc;
  auto c = std::make_shared<Tensor>();
  int g = 0;
  for (auto *list : {&con, &att})
    for (size_t m : *list) {
      ++g;
      kind.push_back(list == &con ? LNT_UNARY_CONTRACT : LNT_UNARY_ATTRACT);
      for (size_t q = 0; q < n; ++q)
        if (la[q].id == la[m].id) group[q] = g;
    }
  for (size_t m = 0; m < n; ++m)
    if (!group[m]) c->dims.push_back(a.dims[m]), lc.push_back(la[m]);
  for (size_t m : att) c->dims.push_back(a.dims[m]), lc.push_back(la[m]);
  c->seq = identity(c->dims.size());
  c->sorted = true;
  int64_t *ck = nullptr;
  double *cv = nullptr;
  int64_t nc = 0;
  check(lnt_contract_attract(default_context().get(), a.keys(), a.vals(), a.nnz(), (int)n, a.dims.data(),
                             a.seq.data(), group.data(), kind.data(), (int)kind.size(), &ck, &cv, &nc, a.mem()));
  c->adopt(ck, cv, nc, a.mem());
  return {c, lc};
}

// Lattice product / Hadamard (MIA_Expr.h:449-561, 864-925, 1012-1059).
inline Expr multiply(Tensor &a, const std::vector<Label> &la, Tensor &b_in, const std::vector<Label> &lb) {
  Tensor b_copy;
  Tensor *bp = &b_in;
  if (&a == &b_in) {
    b_copy = b_in;
    bp = &b_copy;
  }
  Tensor &b = *bp;
  same_memory(a, b);
  std::vector<int32_t> a_out, a_in, a_int, b_in_m, b_int, b_out;
  for (size_t m = 0; m < la.size(); ++m) {
    int r = find_label(lb, la[m].id);
    if (r < 0) {
      if (la[m].inter()) throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "inter index in one operand only");
      a_out.push_back((int32_t)m);
      continue;
    }
    // One id with different element-wise counts may not match
    // (product_rule_different_elem_wise, Index.h:95-101).
    if (la[m].bangs != lb[r].bangs)
      throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "an index has different element-wise counts on the two sides");
    if (a.dims[m] != b.dims[r]) throw DimensionMismatchException(LNT_ERR_DIMENSION_MISMATCH, "matched dims differ");
    if (la[m].inter()) {
      a_int.push_back((int32_t)m);
      b_int.push_back(r);
    } else {
      a_in.push_back((int32_t)m);
      b_in_m.push_back(r);
    }
  }
  for (size_t r = 0; r < lb.size(); ++r)
    if (find_label(la, lb[r].id) < 0) {
      if (lb[r].inter()) throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "inter index in one operand only");
      b_out.push_back((int32_t)r);
    }
  std::vector<int32_t> seq_a, seq_b;
  for (auto v : {&a_out, &a_in, &a_int}) seq_a.insert(seq_a.end(), v->begin(), v->end());
  for (auto v : {&b_in_m, &b_out, &b_int}) seq_b.insert(seq_b.end(), v->begin(), v->end());
  auto dprod = [](const Tensor &t, const std::vector<int32_t> &ms) {
    int64_t p = 1;
    for (int32_t m : ms) p *= t.dims[m];
    return p;
  };
  const int64_t lat_a[3] = {dprod(a, a_out), dprod(a, a_in), dprod(a, a_int)};
  const int64_t lat_b[3] = {lat_a[1], dprod(b, b_out), lat_a[2]};
  auto c = std::make_shared<Tensor>();
  std::vector<Label> lc;
  for (int32_t m : a_out) c->dims.push_back(a.dims[m]), lc.push_back(la[m]);
  for (int32_t r : b_out) c->dims.push_back(b.dims[r]), lc.push_back(lb[r]);
  for (int32_t m : a_int) c->dims.push_back(a.dims[m]), lc.push_back(la[m]);
  // Every index contracted: a 1 x 1 x 1 lattice product, the scalar of the
  // reference's complete inner product (sparse_mia_mult_test.cpp:159-165).
  c->seq = identity(c->dims.size());
  c->sorted = true;
  int64_t *ck = nullptr;
  double *cv = nullptr;
  int64_t nc = 0;
  if (a_in.empty() && a_out.empty() && b_out.empty()) {
    // Pure Hadamard: one sorted merge in any common order.  Every mode is an
    // inter index (a_int = A's modes in order, C's modes = A's), so a sorted
    // A keeps its order, B follows it and C comes out in it.
    if (a.sorted) {
      for (size_t t = 0; t < seq_a.size(); ++t) {
        seq_a[t] = a.seq[t];
        seq_b[t] = b_int[a.seq[t]];
        c->seq[t] = a.seq[t];
      }
    }
    a.sort_to(seq_a);
    b.sort_to(seq_b);
    check(lnt_merge(default_context().get(), a.keys(), a.vals(), a.nnz(), b.keys(), b.vals(), b.nnz(),
                    LNT_OP_HADAMARD, &ck, &cv, &nc, a.mem()));
    c->adopt(ck, cv, nc, a.mem());
    return {c, lc, a_int.size()};
  }
  a.sort_to(seq_a);
  b.sort_to(seq_b);
  check(lnt_lattice_mult(default_context().get(), a.keys(), a.vals(), a.nnz(), lat_a, b.keys(), b.vals(), b.nnz(),
                         lat_b, LNT_SPGEMM_AUTO, &ck, &cv, &nc, a.mem()));
  c->adopt(ck, cv, nc, a.mem());
  a.sorted = false;  // lattice unfolding leaves operands flagged unsorted
  b.sorted = false;
  return {c, lc, a_int.size()};
}

// Dense operand of a product: dims and elements, first mode fastest.
struct DenseView {
  const std::vector<int64_t> *dims;
  const std::vector<double> *data;
};

// Sparse x dense with no contracted index (SparseMIABase::noLatticeMult with
// a DenseMIA, SparseMIABase.h:1829-1910): A sorted to [outer, inter] order,
// then C(l, r, t) = A(l, t) * B(r, t) for non-zero factors.  C's modes are
// [A outer, B outer, inter]; its keys run B outer fastest, then A outer, then
// inter (:1899).
inline Expr multiply_dense(Tensor &a, const std::vector<Label> &la, const DenseView &b, const std::vector<Label> &lb) {
  std::vector<int32_t> a_out, a_int, b_out, b_int;
  for (size_t m = 0; m < la.size(); ++m) {
    const int r = find_label(lb, la[m].id);
    if (r < 0) {
      if (la[m].inter()) throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "inter index in one operand only");
      a_out.push_back((int32_t)m);
      continue;
    }
    if (!la[m].inter() || la[m].bangs != lb[r].bangs)
      throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT,
                                  "a sparse x dense product takes only element-wise shared indices");
    if (a.dims[m] != (*b.dims)[r]) throw DimensionMismatchException(LNT_ERR_DIMENSION_MISMATCH, "matched dims differ");
    a_int.push_back((int32_t)m);
    b_int.push_back(r);
  }
  for (size_t r = 0; r < lb.size(); ++r)
    if (find_label(la, lb[r].id) < 0) {
      if (lb[r].inter()) throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "inter index in one operand only");
      b_out.push_back((int32_t)r);
    }
  std::vector<int32_t> seq_a(a_out);
  seq_a.insert(seq_a.end(), a_int.begin(), a_int.end());
  a.sort_to(seq_a);
  int64_t lsize = 1;
  for (int32_t m : a_out) lsize *= a.dims[m];
  auto c = std::make_shared<Tensor>();
  std::vector<Label> lc;
  for (int32_t m : a_out) c->dims.push_back(a.dims[m]), lc.push_back(la[m]);
  for (int32_t r : b_out) c->dims.push_back((*b.dims)[r]), lc.push_back(lb[r]);
  for (int32_t m : a_int) c->dims.push_back(a.dims[m]), lc.push_back(la[m]);
  const int nl = (int)a_out.size(), nr = (int)b_out.size(), ni = (int)a_int.size();
  for (int q = 0; q < nr; ++q) c->seq.push_back(nl + q);
  for (int q = 0; q < nl; ++q) c->seq.push_back(q);
  for (int q = 0; q < ni; ++q) c->seq.push_back(nl + nr + q);
  c->sorted = true;
  int64_t *ck = nullptr;
  double *cv = nullptr;
  int64_t nc = 0;
  // A dense operand meeting a sparse one in HBM goes there for the call.
  std::shared_ptr<DeviceArray> bd;
  const double *bdata = b.data->data();
  if (a.device) {
    bd = DeviceArray::copy_of(bdata, (int64_t)b.data->size() * 8);
    bdata = (const double *)bd->p;
  }
  check(lnt_sparse_dense_mult(default_context().get(), a.keys(), a.vals(), a.nnz(), lsize, (int)b.dims->size(),
                              b.dims->data(), bdata, b_out.data(), nr, b_int.data(), ni, &ck, &cv, &nc, a.mem()));
  c->adopt(ck, cv, nc, a.mem());
  return {c, lc, (size_t)ni};
}

}  // namespace detail

// B(i,j,...) for a DenseMIA B: the right operand of a sparse x dense product.
struct DenseAtom {
  detail::DenseView view;
  std::vector<Label> labels;
};

// Expression node: A(i,j,...) or the result of an operation on atoms.
class Atom {
 public:
  Atom(detail::Expr e) : e_(std::move(e)) {}
  Atom operator+(const Atom &o) const { return detail::merge(*e_.t, e_.labels, *o.e_.t, o.e_.labels, LNT_OP_ADD); }
  Atom operator-(const Atom &o) const { return detail::merge(*e_.t, e_.labels, *o.e_.t, o.e_.labels, LNT_OP_SUB); }
  Atom operator*(const Atom &o) const { return detail::multiply(*e_.t, e_.labels, *o.e_.t, o.e_.labels); }
  Atom operator*(const DenseAtom &o) const { return detail::multiply_dense(*e_.t, e_.labels, o.view, o.labels); }
  // ~(A*B): the product's inter indices lose one `!` (MIA_Expr.h:257-301,
  // 1376), so `!j` becomes a plain index of the next product and `!!l` an
  // inter index again.  No data moves.
  Atom operator~() const {
    detail::Expr e = e_;
    for (size_t q = e.labels.size() - e.inter_count; q < e.labels.size(); ++q)
      e.labels[q].bangs = std::max(e.labels[q].bangs - 1, 0);
    e.inter_count = 0;
    return Atom(e);
  }
  // The value of a complete inner product (every index contracted).
  operator double() const {
    if (!e_.labels.empty())
      throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "only a complete inner product is a scalar");
    detail::Tensor &t = *e_.t;
    if (t.nnz() == 0) return 0.0;
    if (!t.device) return t.data[0];
    double v = 0.0;
    check(lnt_copy(default_context().get(), &v, t.vals(), sizeof(double)));
    return v;
  }
  const detail::Expr &expr() const { return e_; }

 private:
  detail::Expr e_;
};

// Left-hand side of an assignment: C(i,j,...) = expr (SparseMIA::assign,
// SparseMIA.h:891-928 -- relabels the mode order, moves no index data).
template <class MIA>
class AssignTarget {
 public:
  AssignTarget(MIA &m, std::vector<Label> labels) : m_(m), labels_(std::move(labels)) {}
  // A repeated index contracts or attracts at once (perform_unary).
  operator Atom() const { return Atom(detail::unary(m_.tensor_ptr(), labels_)); }
  Atom operator+(const Atom &o) const { return Atom(*this) + o; }
  Atom operator-(const Atom &o) const { return Atom(*this) - o; }
  Atom operator*(const Atom &o) const { return Atom(*this) * o; }
  Atom operator+(const AssignTarget &o) const { return Atom(*this) + Atom(o); }
  Atom operator-(const AssignTarget &o) const { return Atom(*this) - Atom(o); }
  Atom operator*(const AssignTarget &o) const { return Atom(*this) * Atom(o); }
  Atom operator*(const DenseAtom &o) const { return Atom(*this) * o; }
  // Assignment matches indices by id, whatever their element-wise counts
  // (MIA_Expr.h:112-135).
  AssignTarget &operator=(const Atom &rhs) {
    for (size_t m = 0; m < labels_.size(); ++m)
      if (detail::find_label(labels_, labels_[m].id) != (int)m)
        throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "an assignment target repeats an index");
    const detail::Expr &e = rhs.expr();
    if (e.labels.size() != labels_.size())
      throw DimensionMismatchException(LNT_ERR_DIMENSION_MISMATCH, "assignment order differs");
    const size_t n = labels_.size();
    std::vector<int32_t> perm(n), inv(n);  // C mode m <- result mode perm[m]
    for (size_t m = 0; m < n; ++m) {
      int p = detail::find_label(e.labels, labels_[m].id);
      if (p < 0) throw MIAParameterException(LNT_ERR_INVALID_ARGUMENT, "assignment indices differ");
      perm[m] = p;
      inv[p] = (int32_t)m;
    }
    detail::Tensor &t = *m_.tensor_ptr();
    // A temporary result is moved, a tensor still referenced elsewhere copied.
    detail::Tensor src = e.t.use_count() == 1 ? std::move(*e.t) : detail::Tensor(*e.t);
    std::vector<int64_t> dims(n);
    for (size_t m = 0; m < n; ++m) dims[m] = src.dims[perm[m]];
    if (dims != t.dims) throw DimensionMismatchException(LNT_ERR_DIMENSION_MISMATCH, "assignment dims differ");
    std::vector<int32_t> seq(n);
    for (size_t q = 0; q < n; ++q) seq[q] = inv[src.seq[q]];
    t.idx = std::move(src.idx);
    t.data = std::move(src.data);
    t.device = src.device;
    t.dk = std::move(src.dk);
    t.dv = std::move(src.dv);
    t.dn = src.dn;
    t.seq = seq;
    t.sorted = src.sorted;
    return *this;
  }
  AssignTarget &operator=(const AssignTarget &rhs) { return *this = Atom(rhs); }
  // Destructive merge A(...) += B(...) / -= (SparseMIA::merge,
  // SparseMIA.h:1030-1050: a scan merge when either operand is sorted, a sort
  // merge otherwise) -- the merge of this target and the right-hand side,
  // assigned back to this target.
  AssignTarget &operator+=(const Atom &rhs) { return *this = (Atom(*this) + rhs); }
  AssignTarget &operator-=(const Atom &rhs) { return *this = (Atom(*this) - rhs); }
  AssignTarget &operator+=(const AssignTarget &rhs) { return *this += Atom(rhs); }
  AssignTarget &operator-=(const AssignTarget &rhs) { return *this -= Atom(rhs); }

 private:
  MIA &m_;
  std::vector<Label> labels_;
};

// ------------------------------------------------------------- SparseMIA --
template <class T, size_t N>
class SparseMIA {
  static_assert(std::is_same<T, double>::value, "the GPU path computes in double");

 public:
  using index_type = int64_t;
  using data_type = T;

  SparseMIA() : t_(std::make_shared<detail::Tensor>()) {
    t_->dims.assign(N, 0);
    t_->seq = detail::identity(N);
  }
  template <class... D, class = typename std::enable_if<sizeof...(D) == N>::type>
  explicit SparseMIA(D... dims) : SparseMIA() {
    t_->dims = {static_cast<int64_t>(dims)...};
  }
  explicit SparseMIA(const std::array<int64_t, N> &dims) : SparseMIA() { t_->dims.assign(dims.begin(), dims.end()); }
  SparseMIA(const SparseMIA &o) : t_(std::make_shared<detail::Tensor>(*o.t_)) {}
  SparseMIA &operator=(const SparseMIA &o) {
    *t_ = *o.t_;
    return *this;
  }

  // ---- storage (SparseMIA.h:61-65, 203-206)
  size_t size() const { return (size_t)t_->nnz(); }
  std::array<int64_t, N> dims() const {
    std::array<int64_t, N> d;
    std::copy(t_->dims.begin(), t_->dims.end(), d.begin());
    return d;
  }
  int64_t dim(size_t i) const { return t_->dims[i]; }
  int64_t dimensionality() const { return detail::product(t_->dims); }
  std::array<size_t, N> linIdxSequence() const {
    std::array<size_t, N> s;
    for (size_t i = 0; i < N; ++i) s[i] = (size_t)t_->seq[i];
    return s;
  }
  bool is_sorted() const { return t_->sorted; }
  void setSorted(bool s) { t_->sorted = s; }
  void push_back(T value, int64_t index) {
    host_only();
    t_->data.push_back(value);
    t_->idx.push_back(index);
    t_->sorted = false;
  }
  void reserve(size_t n) {
    host_only();
    t_->idx.reserve(n);
    t_->data.reserve(n);
  }
  void resize(size_t n) {
    host_only();
    t_->idx.resize(n);
    t_->data.resize(n);
  }
  int64_t *index_begin() { return host_only(), t_->idx.data(); }
  int64_t *index_end() { return host_only(), t_->idx.data() + t_->idx.size(); }
  T *data_begin() { return host_only(), t_->data.data(); }
  T *data_end() { return host_only(), t_->data.data() + t_->data.size(); }
  const std::vector<int64_t> &indices() const { return host_only(), t_->idx; }
  const std::vector<T> &data() const { return host_only(), t_->data; }

  // ---- residence in HBM (no reference counterpart: its storage is host
  // memory).  While on the device the entries are reached only through
  // operations; the host accessors above throw.
  void to_device() { t_->to_device(); }
  void to_host() { t_->to_host(); }
  bool on_device() const { return t_->device; }
  // Device pointers to the n entries in HBM, valid while the tensor stays
  // on the device and unchanged.
  const int64_t *device_indices() const { return t_->device ? (const int64_t *)t_->dk->p : nullptr; }
  const T *device_data() const { return t_->device ? (const T *)t_->dv->p : nullptr; }
  // Takes over n entries already in HBM: keys and vals allocated with
  // lnt_alloc (LNT_MEM_DEVICE), in the mode order seq, sorted or not; the
  // tensor frees them.
  template <class S>
  void assign_device(int64_t *keys, double *vals, int64_t n, const std::array<S, N> &seq, bool sorted) {
    t_->adopt(keys, vals, n, LNT_MEM_DEVICE);
    t_->seq.assign(seq.begin(), seq.end());
    t_->sorted = sorted;
  }

  // ---- linear-index math (IndexUtil.h:318-333, 866-879)
  int64_t sub2ind(const std::array<int64_t, N> &sub) const {
    int64_t k = 0, mult = 1;
    for (size_t t = 0; t < N; ++t) {
      k += sub[t_->seq[t]] * mult;
      mult *= t_->dims[t_->seq[t]];
    }
    return k;
  }
  std::array<int64_t, N> ind2sub(int64_t k) const {
    std::array<int64_t, N> sub{};
    for (size_t t = 0; t < N; ++t) {
      sub[t_->seq[t]] = k % t_->dims[t_->seq[t]];
      k /= t_->dims[t_->seq[t]];
    }
    return sub;
  }

  // ---- sorts (SparseMIABase.h:590-645, 701-734, 768-814)
  void sort() { sort(linIdxSequence()); }
  template <class S>
  void sort(const std::array<S, N> &seq) {
    std::vector<int32_t> s(seq.begin(), seq.end());
    t_->sort_to(s);
  }
  template <class S>
  void change_linIdx_sequence(const std::array<S, N> &seq) {
    std::vector<int32_t> s(seq.begin(), seq.end());
    if (s == t_->seq || N == 1) return;
    t_->relinearise(s);
  }

  // ---- comparison (SparseMIABase.h:1243-1296, 1616-1635): dims must agree;
  // this is sorted, the other sorted into this one's sequence (both in
  // place), then compared on the GPU; entries with |value| <= 1.11e-16 count
  // as absent.
  bool operator==(SparseMIA &o) { return compare(o, LNT_CMP_EXACT, 0.0); }
  bool operator!=(SparseMIA &o) { return !(*this == o); }
  bool fuzzy_equals(SparseMIA &o, T precision) { return compare(o, LNT_CMP_FUZZY, precision); }

  // ---- index expressions (MIA.h:222-229)
  template <class... Ts>
  AssignTarget<SparseMIA> operator()(Ts...) {
    static_assert(sizeof...(Ts) == N, "one index per mode");
    return AssignTarget<SparseMIA>(*this, std::vector<Label>{Label{Ts::id, Ts::elemval}...});
  }

  std::shared_ptr<detail::Tensor> tensor_ptr() { return t_; }

 private:
  bool compare(SparseMIA &o, int mode, double precision) {
    if (t_->dims != o.t_->dims) return false;
    detail::same_memory(*t_, *o.t_);
    sort();
    o.t_->sort_to(t_->seq);
    int eq = 0;
    check(lnt_compare(default_context().get(), t_->keys(), t_->vals(), t_->nnz(), o.t_->keys(), o.t_->vals(),
                      o.t_->nnz(), mode, precision, &eq, t_->mem()));
    return eq != 0;
  }
  void host_only() const {
    if (t_->device)
      throw MIAException(LNT_ERR_INVALID_ARGUMENT, "the entries are in HBM: call to_host() to reach them");
  }

  std::shared_ptr<detail::Tensor> t_;
};

// ------------------------------------------------------------- DenseMIA --
// Dense order-N tensor, elements first mode fastest (the reference's
// DenseMIA layout, multi_index_array/DenseMIA.h): the dense operand of
// sparse x dense products A(l,!t) * B(r,!t).
template <class T, size_t N>
class DenseMIA {
  static_assert(std::is_same<T, double>::value, "the GPU path computes in double");

 public:
  template <class... D, class = typename std::enable_if<sizeof...(D) == N>::type>
  explicit DenseMIA(D... dims) : dims_{static_cast<int64_t>(dims)...} {
    data_.assign((size_t)detail::product(dims_), 0.0);
  }
  int64_t dim(size_t i) const { return dims_[i]; }
  int64_t dimensionality() const { return detail::product(dims_); }
  T *data_begin() { return data_.data(); }
  T *data_end() { return data_.data() + data_.size(); }
  const std::vector<T> &data() const { return data_; }
  // Element at full subscript, first mode fastest.
  T &at(const std::array<int64_t, N> &sub) {
    int64_t k = 0, mult = 1;
    for (size_t t = 0; t < N; ++t) k += sub[t] * mult, mult *= dims_[t];
    return data_[(size_t)k];
  }
  template <class... Ts>
  DenseAtom operator()(Ts...) const {
    static_assert(sizeof...(Ts) == N, "one index per mode");
    return DenseAtom{detail::DenseView{&dims_, &data_}, std::vector<Label>{Label{Ts::id, Ts::elemval}...}};
  }

 private:
  std::vector<int64_t> dims_;
  std::vector<T> data_;
};

}  // namespace lnt

#define LNT_UNIQUE_ID __COUNTER__
#define NTINDEX ::lnt::ProdInd<LNT_UNIQUE_ID>
#define MIAINDEX ::lnt::ProdInd<LNT_UNIQUE_ID>
