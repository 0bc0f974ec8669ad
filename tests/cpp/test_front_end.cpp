// C++ front-end test in the style of the reference's own tests
// (src/tests/SparseMIA/sparse_mia_functions.cpp:100-202,
// sparse_mia_add_sub_tests.cpp, sparse_mia_mult_test.cpp): every sparse result
// is compared with dense brute force computed here on the CPU.
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "lnt/SparseMIA.hpp"

static int failures = 0;
#define CHECK(cond, msg)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      std::printf("FAIL: %s (line %d)\n", msg, __LINE__); \
      ++failures;                                         \
    }                                                     \
  } while (0)

template <size_t N>
static void fill_random(lnt::SparseMIA<double, N> &m, size_t nnz, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::map<int64_t, double> entries;
  std::uniform_int_distribution<int64_t> key(0, m.dimensionality() - 1);
  std::uniform_real_distribution<double> val(-1.0, 1.0);
  while (entries.size() < nnz) entries[key(rng)] = val(rng);
  std::vector<std::pair<int64_t, double>> v(entries.begin(), entries.end());
  std::shuffle(v.begin(), v.end(), rng);
  for (auto &e : v) m.push_back(e.second, e.first);
}

// Dense value at full subscript `sub` (default first-mode-fastest layout).
template <size_t N>
static std::vector<double> dense(const lnt::SparseMIA<double, N> &m) {
  std::vector<double> d(m.dimensionality(), 0.0);
  for (size_t q = 0; q < m.size(); ++q) {
    auto sub = m.ind2sub(m.indices()[q]);
    int64_t k = 0, mult = 1;
    for (size_t t = 0; t < N; ++t) {
      k += sub[t] * mult;
      mult *= m.dim(t);
    }
    d[k] = m.data()[q];
  }
  return d;
}

// Dense value at full subscript `sub` of a dense copy with dims `dims`.
static double at(const std::vector<double> &d, std::initializer_list<int64_t> dims,
                 std::initializer_list<int64_t> sub) {
  int64_t k = 0, mult = 1;
  auto di = dims.begin();
  for (int64_t s : sub) {
    k += s * mult;
    mult *= *di++;
  }
  return d[k];
}

static bool near(double got, double want) { return std::fabs(got - want) <= 1e-12 * std::max(1.0, std::fabs(want)); }

// Dense einsum-like reference: `in` labels one letter per mode of a dense
// array with dims `dims` (first mode fastest); `out` lists the result's
// letters.  Letters repeated in `in` must agree; the rest are summed.
static std::vector<double> dense_unary(const std::vector<double> &d, const std::vector<int64_t> &dims,
                                       const std::string &in, const std::string &out) {
  std::vector<int64_t> odims;
  for (char c : out) odims.push_back(dims[in.find(c)]);
  int64_t osize = 1;
  for (int64_t x : odims) osize *= x;
  std::vector<double> r(osize, 0.0);
  for (size_t q = 0; q < d.size(); ++q) {
    std::vector<int64_t> sub(dims.size());
    int64_t k = (int64_t)q;
    for (size_t t = 0; t < dims.size(); ++t) sub[t] = k % dims[t], k /= dims[t];
    bool ok = true;
    for (size_t t = 0; t < in.size(); ++t) ok = ok && sub[t] == sub[in.find(in[t])];
    if (!ok) continue;
    int64_t o = 0, mult = 1;
    for (size_t u = 0; u < out.size(); ++u) o += sub[in.find(out[u])] * mult, mult *= odims[u];
    r[o] += d[q];
  }
  return r;
}

int main() {
  NTINDEX i; NTINDEX j; NTINDEX k; NTINDEX l; NTINDEX m; NTINDEX n;
  // ---- sort(seq) == change_linIdx_sequence(seq) + sort()  (sparse_mia_functions.cpp:100-202)
  {
    lnt::SparseMIA<double, 4> a(7, 9, 5, 9), b(7, 9, 5, 9);
    fill_random(a, 1200, 1);
    b = a;
    std::array<size_t, 4> seq{{2, 0, 3, 1}};
    a.sort(seq);
    b.change_linIdx_sequence(seq);
    b.sort();
    CHECK(a.indices() == b.indices() && a.data() == b.data(), "sort(seq) differs from relinearise+sort");
    for (size_t q = 1; q < a.size(); ++q) CHECK(a.indices()[q - 1] < a.indices()[q], "not ascending");
  }
  // ---- C(i,j,k,l) = A(i,j,k,l) +/- B(k,l,i,j)  (sparse_mia_add_sub_tests.cpp)
  {
    lnt::SparseMIA<double, 4> a(6, 5, 4, 3), b(4, 3, 6, 5), c(6, 5, 4, 3), c2(6, 5, 4, 3);
    fill_random(a, 150, 2);
    fill_random(b, 200, 3);
    auto da = dense(a), db = dense(b);
    c(i, j, k, l) = a(i, j, k, l) + b(k, l, i, j);
    c2(i, j, k, l) = a(i, j, k, l) - b(k, l, i, j);
    auto dc = dense(c), dc2 = dense(c2);
    for (int64_t q = 0; q < c.dimensionality(); ++q) {
      int64_t ii = q % 6, jj = (q / 6) % 5, kk = (q / 30) % 4, ll = q / 120;
      double bv = db[kk + 4 * ll + 12 * ii + 72 * jj];
      CHECK(dc[q] == da[q] + bv, "A+B mismatch");
      CHECK(dc2[q] == da[q] - bv, "A-B mismatch");
    }
  }
  // ---- neither operand sorted: sortMerge, whose collect_duplicates drops
  // keys whose entries are all zero (SparseMIA.h:628-665, 833-887)
  {
    lnt::SparseMIA<double, 2> a(4, 4), b(4, 4), c(4, 4);
    a.push_back(1.0, 5);
    a.push_back(0.0, 9);
    a.push_back(2.0, 3);
    b.push_back(-2.0, 3);
    b.push_back(0.0, 11);
    b.push_back(4.0, 1);
    c(i, j) = a(i, j) + b(i, j);
    CHECK(c.size() == 3, "sortMerge result size");
    if (c.size() == 3) {
      CHECK(c.indices()[0] == 1 && c.indices()[1] == 3 && c.indices()[2] == 5, "sortMerge keys");
      CHECK(c.data()[0] == 4.0 && c.data()[1] == 0.0 && c.data()[2] == 1.0, "sortMerge values");
    }
  }
  // ---- C(i,j,m,n) = A(i,j,k,l) * B(k,l,m,n)  (sparse_mia_mult_test.cpp)
  {
    lnt::SparseMIA<double, 4> a(5, 4, 6, 3), b(6, 3, 4, 5), c(5, 4, 4, 5);
    fill_random(a, 120, 4);
    fill_random(b, 110, 5);
    auto da = dense(a), db = dense(b);
    c(i, j, m, n) = a(i, j, k, l) * b(k, l, m, n);
    auto dc = dense(c);
    double max_err = 0;
    for (int ii = 0; ii < 5; ++ii)
      for (int jj = 0; jj < 4; ++jj)
        for (int mm = 0; mm < 4; ++mm)
          for (int nn = 0; nn < 5; ++nn) {
            double s = 0;
            for (int kk = 0; kk < 6; ++kk)
              for (int ll = 0; ll < 3; ++ll) s += da[ii + 5 * jj + 20 * kk + 120 * ll] * db[kk + 6 * ll + 18 * mm + 72 * nn];
            max_err = std::max(max_err, std::fabs(s - dc[ii + 5 * jj + 20 * mm + 80 * nn]));
          }
    CHECK(max_err < 1e-12, "contraction mismatch");
  }
  // ---- Hadamard C(i,j) = A(!i,!j) * B(!j,!i)
  {
    lnt::SparseMIA<double, 2> a(30, 20), b(20, 30), c(30, 20);
    fill_random(a, 300, 6);
    fill_random(b, 300, 7);
    auto da = dense(a), db = dense(b);
    c(i, j) = a(!i, !j) * b(!j, !i);
    auto dc = dense(c);
    for (int ii = 0; ii < 30; ++ii)
      for (int jj = 0; jj < 20; ++jj)
        CHECK(dc[ii + 30 * jj] == da[ii + 30 * jj] * db[jj + 20 * ii], "Hadamard mismatch");
  }
  // ---- errors become exceptions (LibMIAException.h)
  {
    lnt::SparseMIA<double, 2> a(3, 4), b(5, 4), c(3, 4);
    fill_random(a, 5, 8);
    fill_random(b, 5, 9);
    bool thrown = false;
    try {
      c(i, j) = a(i, j) + b(i, j);
    } catch (const lnt::DimensionMismatchException &) {
      thrown = true;
    }
    CHECK(thrown, "dimension mismatch not reported");
  }
  // ---- destructive merges A(i,j,k) += / -= B(...) (SparseMIA.h:1030-1050)
  {
    lnt::SparseMIA<double, 3> a(5, 6, 7), b(7, 5, 6), a0(5, 6, 7);
    fill_random(a, 60, 31);
    fill_random(b, 70, 32);
    a0 = a;
    std::vector<double> da = dense(a0), db = dense(b);
    a(i, j, k) += b(k, i, j);
    std::vector<double> dc = dense(a);
    for (int64_t ii = 0; ii < 5; ++ii)
      for (int64_t jj = 0; jj < 6; ++jj)
        for (int64_t kk = 0; kk < 7; ++kk)
          CHECK(dc[ii + 5 * (jj + 6 * kk)] == da[ii + 5 * (jj + 6 * kk)] + db[kk + 7 * (ii + 5 * jj)], "A += B mismatch");
    a(i, j, k) -= b(k, i, j);  // per key: (a + b) - b, or a where b has no entry
    std::vector<double> dd = dense(a);
    for (int64_t ii = 0; ii < 5; ++ii)
      for (int64_t jj = 0; jj < 6; ++jj)
        for (int64_t kk = 0; kk < 7; ++kk) {
          const int64_t q = ii + 5 * (jj + 6 * kk);
          CHECK(dd[q] == dc[q] - db[kk + 7 * (ii + 5 * jj)], "A -= B mismatch");
        }
    lnt::SparseMIA<double, 3> z(5, 6, 7);
    z = a0;
    z(i, j, k) -= a0(i, j, k);
    for (double v : dense(z)) CHECK(v == 0.0, "A -= A is not zero");
  }
  // ---- comparison (SparseMIABase.h:1243-1296, 1616-1635)
  {
    lnt::SparseMIA<double, 3> a(6, 7, 8), b(6, 7, 8);
    fill_random(a, 100, 21);
    b = a;
    b.sort(std::array<size_t, 3>{{2, 0, 1}});  // same tensor, another order
    CHECK(a == b, "a tensor differs from itself in another order");
    int64_t absent = 0;  // a key that a does not hold
    while (std::find(a.indices().begin(), a.indices().end(), absent) != a.indices().end()) ++absent;
    b.change_linIdx_sequence(a.linIdxSequence());
    b.push_back(1e-17, absent);  // below the sparse tolerance: counts as absent
    CHECK(a == b, "an entry below the tolerance broke equality");
    lnt::SparseMIA<double, 3> c(6, 7, 8);
    c = a;
    c.data_begin()[3] *= 1.0 + 1e-9;
    c.setSorted(false);
    CHECK(!(a == c), "a changed value compared equal");
    CHECK(a.fuzzy_equals(c, 1e-6), "fuzzy_equals rejected a tiny change");
    lnt::SparseMIA<double, 3> d(6, 7, 9);
    CHECK(!(a == d), "different dims compared equal");
  }
  // ---- chained products with ~ and repeated element-wise indices, complete
  //      inner products (sparse_mia_mult_test.cpp:135-165)
  {
    const int64_t I = 3, J = 4, K = 5, L = 6, M = 3, N = 2;
    lnt::SparseMIA<double, 4> a(I, J, K, L), b(J, L, M, N), c(I, J, K, L), ct(I, K, M, N);
    lnt::SparseMIA<double, 2> b2(J, L), d2(J, L), d2t(L, J), sq(J, J);
    lnt::SparseMIA<double, 1> d(L);
    lnt::SparseMIA<double, 3> c3(J, J, J);
    fill_random(a, 250, 31);
    fill_random(b, 100, 32);
    fill_random(b2, 15, 33);
    fill_random(d2, 14, 34);
    fill_random(d2t, 13, 35);
    fill_random(d, 5, 36);
    fill_random(sq, 9, 37);
    auto da = dense(a), db = dense(b), db2 = dense(b2), dd = dense(d), dd2 = dense(d2), dd2t = dense(d2t);
    auto dsq = dense(sq);
    // Repeated Element-Wise Product 1: c(i,j,k,l) = ~(a(i,!j,k,!!l)*b2(!j,!!l))*d(!l)
    c(i, j, k, l) = ~(a(i, !j, k, !!l) * b2(!j, !!l)) * d(!l);
    auto dc = dense(c);
    for (int64_t ii = 0; ii < I; ++ii)
      for (int64_t jj = 0; jj < J; ++jj)
        for (int64_t kk = 0; kk < K; ++kk)
          for (int64_t ll = 0; ll < L; ++ll) {
            double want = at(da, {I, J, K, L}, {ii, jj, kk, ll}) * at(db2, {J, L}, {jj, ll}) * dd[ll];
            CHECK(near(at(dc, {I, J, K, L}, {ii, jj, kk, ll}), want), "repeated element-wise product");
          }
    // Ternary Inner Products 1 and 2: ~(a(i,!j,k,!l)*b(!j,!l,m,n))*d2(j,l) / d2(l,j)
    for (int variant = 0; variant < 2; ++variant) {
      if (variant == 0)
        ct(i, k, m, n) = ~(a(i, !j, k, !l) * b(!j, !l, m, n)) * d2(j, l);
      else
        ct(i, k, m, n) = ~(a(i, !j, k, !l) * b(!j, !l, m, n)) * d2t(l, j);
      auto dct = dense(ct);
      for (int64_t ii = 0; ii < I; ++ii)
        for (int64_t kk = 0; kk < K; ++kk)
          for (int64_t mm = 0; mm < M; ++mm)
            for (int64_t nn = 0; nn < N; ++nn) {
              double want = 0.0;
              for (int64_t jj = 0; jj < J; ++jj)
                for (int64_t ll = 0; ll < L; ++ll)
                  want += at(da, {I, J, K, L}, {ii, jj, kk, ll}) * at(db, {J, L, M, N}, {jj, ll, mm, nn}) *
                          (variant == 0 ? at(dd2, {J, L}, {jj, ll}) : at(dd2t, {L, J}, {ll, jj}));
              CHECK(near(at(dct, {I, K, M, N}, {ii, kk, mm, nn}), want), "ternary inner product");
            }
    }
    // Outer/Element-Wise Product with self-multiplication: c3(i,j,l) = sq(i,!j)*sq(!j,l)
    c3(i, j, l) = sq(i, !j) * sq(!j, l);
    auto dc3 = dense(c3);
    for (int64_t ii = 0; ii < J; ++ii)
      for (int64_t jj = 0; jj < J; ++jj)
        for (int64_t ll = 0; ll < J; ++ll)
          CHECK(near(at(dc3, {J, J, J}, {ii, jj, ll}), at(dsq, {J, J}, {ii, jj}) * at(dsq, {J, J}, {jj, ll})),
                "self element-wise product");
    // Complete inner products: a(i,j,k,l)*e(i,l,j,k) and with itself
    lnt::SparseMIA<double, 4> e(I, L, J, K), f(I, J, K, J);
    fill_random(e, 200, 38);
    fill_random(f, 150, 39);
    auto de = dense(e), df = dense(f);
    double s1 = a(i, j, k, l) * e(i, l, j, k), w1 = 0.0;
    double s2 = f(i, l, k, j) * f(i, j, k, l), w2 = 0.0;
    for (int64_t ii = 0; ii < I; ++ii)
      for (int64_t jj = 0; jj < J; ++jj)
        for (int64_t kk = 0; kk < K; ++kk) {
          for (int64_t ll = 0; ll < L; ++ll)
            w1 += at(da, {I, J, K, L}, {ii, jj, kk, ll}) * at(de, {I, L, J, K}, {ii, ll, jj, kk});
          for (int64_t ll = 0; ll < J; ++ll)
            w2 += at(df, {I, J, K, J}, {ii, ll, kk, jj}) * at(df, {I, J, K, J}, {ii, jj, kk, ll});
        }
    CHECK(near(s1, w1), "complete inner product");
    CHECK(near(s2, w2), "complete inner product with itself");
    bool thrown = false;
    try {
      c(i, j, k, l) = a(i, !j, k, l) * b2(j, l);  // one id, different element-wise counts
    } catch (const std::exception &) {
      thrown = true;
    }
    CHECK(thrown, "mismatched element-wise counts not reported");
  }
  // ---- unary contraction and attraction (tests/MIA/mia_unary_test.cpp:52-171)
  {
    lnt::SparseMIA<double, 6> a(5, 6, 5, 6, 5, 3), a2(5, 6, 5, 5, 6, 3);
    lnt::SparseMIA<double, 3> c(5, 6, 5);
    lnt::SparseMIA<double, 2> sq(5, 5);
    fill_random(a, 1500, 41);
    fill_random(a2, 1400, 42);
    fill_random(c, 50, 43);
    fill_random(sq, 10, 44);
    a.sort(std::array<size_t, 6>{{3, 1, 5, 0, 2, 4}});  // any mode order
    auto da = dense(a), da2 = dense(a2), dc = dense(c), dsq = dense(sq);
    const std::vector<int64_t> A{5, 6, 5, 6, 5, 3}, A2{5, 6, 5, 5, 6, 3}, Cd{5, 6, 5}, S{5, 5};
    auto same = [](const std::vector<double> &got, const std::vector<double> &want) {
      if (got.size() != want.size()) return false;
      for (size_t q = 0; q < got.size(); ++q)
        if (!near(got[q], want[q])) return false;
      return true;
    };
    lnt::SparseMIA<double, 4> b(6, 6, 5, 3);
    b(j, k, l, m) = a(i, j, i, k, l, m);  // Simple Contraction 1
    CHECK(same(dense(b), dense_unary(da, A, "ijiklm", "jklm")), "simple contraction");
    lnt::SparseMIA<double, 1> e(6);
    e(i) = c(j, i, j);  // Simple Contraction 2
    CHECK(same(dense(e), dense_unary(dc, Cd, "jij", "i")), "simple contraction 2");
    lnt::SparseMIA<double, 2> d(5, 3);
    d(l, m) = a(i, j, i, j, l, m);  // Complex Contraction 1
    CHECK(same(dense(d), dense_unary(da, A, "ijijlm", "lm")), "complex contraction");
    lnt::SparseMIA<double, 1> e2(3);
    e2(m) = a(i, j, i, j, i, m);  // Complex Contraction 2
    CHECK(same(dense(e2), dense_unary(da, A, "ijijim", "m")), "complex contraction 2");
    lnt::SparseMIA<double, 1> e3(5);
    e3(i) = sq(!i, !i);  // Very Simple Attraction
    CHECK(same(dense(e3), dense_unary(dsq, S, "ii", "i")), "simple attraction");
    lnt::SparseMIA<double, 2> d2(5, 6);
    d2(i, j) = c(!i, j, !i);  // Simple Attraction 1
    CHECK(same(dense(d2), dense_unary(dc, Cd, "iji", "ij")), "simple attraction 1");
    lnt::SparseMIA<double, 3> c3(5, 5, 3);
    c3(i, k, m) = a(!i, j, !i, j, k, m);  // Combined Contraction/Attraction 1
    CHECK(same(dense(c3), dense_unary(da, A, "ijijkm", "ikm")), "combined contraction/attraction");
    lnt::SparseMIA<double, 2> d3(6, 3);
    d3(j, m) = a2(i, !j, i, i, !j, m);  // Combined Contraction/Attraction 4
    CHECK(same(dense(d3), dense_unary(da2, A2, "ijiijm", "jm")), "combined contraction/attraction 4");
    // a contraction feeding a product: C(j,n) = A(i,j,i,k,l,m) * B(k,l,m,n)
    lnt::SparseMIA<double, 4> bb(6, 5, 3, 4);
    lnt::SparseMIA<double, 2> cj(6, 4);
    fill_random(bb, 200, 45);
    cj(j, n) = a(i, j, i, k, l, m) * bb(k, l, m, n);
    auto tr = dense_unary(da, A, "ijiklm", "jklm");
    auto dbb = dense(bb), dcj = dense(cj);
    bool ok = true;
    for (int64_t jj = 0; jj < 6; ++jj)
      for (int64_t nn = 0; nn < 4; ++nn) {
        double w = 0;
        for (int64_t q = 0; q < 6 * 5 * 3; ++q) w += tr[jj + 6 * q] * dbb[q + 90 * nn];
        ok = ok && near(dcj[jj + 6 * nn], w);
      }
    CHECK(ok, "contraction feeding a product");
    bool thrown = false;
    try {
      lnt::SparseMIA<double, 2> z(5, 5);
      z(i, i) = sq(i, j);
    } catch (const std::exception &) {
      thrown = true;
    }
    CHECK(thrown, "a repeated index in an assignment target not reported");
  }
  // ---- sparse x dense with no contracted index (SparseMIABase.h:1829-1910)
  {
    lnt::SparseMIA<double, 3> a(4, 5, 6);
    fill_random(a, 60, 51);
    auto da = dense(a);
    lnt::DenseMIA<double, 2> bd(6, 3), be(6, 5);
    std::mt19937_64 rng(52);
    std::uniform_real_distribution<double> val(-1.0, 1.0);
    for (double *p = bd.data_begin(); p != bd.data_end(); ++p) *p = (rng() % 4 == 0) ? 0.0 : val(rng);
    for (double *p = be.data_begin(); p != be.data_end(); ++p) *p = (rng() % 4 == 0) ? 0.0 : val(rng);
    // C(i,j,k,l) = A(i,j,!k) * B(!k,l): outer x element-wise
    lnt::SparseMIA<double, 4> c(4, 5, 6, 3);
    c(i, j, k, l) = a(i, j, !k) * bd(!k, l);
    auto dc = dense(c);
    bool ok = true;
    size_t expect_nnz = 0;
    for (int64_t ii = 0; ii < 4; ++ii)
      for (int64_t jj = 0; jj < 5; ++jj)
        for (int64_t kk = 0; kk < 6; ++kk)
          for (int64_t ll = 0; ll < 3; ++ll) {
            const double w = da[ii + 4 * (jj + 5 * kk)] * bd.data()[kk + 6 * ll];
            ok = ok && dc[ii + 4 * (jj + 5 * (kk + 6 * ll))] == w;
            expect_nnz += w != 0.0;
          }
    CHECK(ok, "sparse x dense (outer x element-wise) mismatch");
    CHECK(c.size() == expect_nnz, "sparse x dense kept a zero factor");
    // C(i,k,j) = A(i,!j,!k) * B(!k,!j): pure element-wise over two indices, B's modes swapped
    lnt::SparseMIA<double, 3> c2(4, 6, 5);
    c2(i, k, j) = a(i, !j, !k) * be(!k, !j);
    auto dc2 = dense(c2);
    ok = true;
    for (int64_t ii = 0; ii < 4; ++ii)
      for (int64_t jj = 0; jj < 5; ++jj)
        for (int64_t kk = 0; kk < 6; ++kk)
          ok = ok && dc2[ii + 4 * (kk + 6 * jj)] == da[ii + 4 * (jj + 5 * kk)] * be.data()[kk + 6 * jj];
    CHECK(ok, "sparse x dense (element-wise) mismatch");
    bool thrown = false;
    try {
      c2(i, k, j) = a(i, j, k) * bd(k, l);  // a contracted index
    } catch (const std::exception &) {
      thrown = true;
    }
    CHECK(thrown, "a contraction with a dense operand not reported");
  }
  // ---- tensors resident in HBM: the same expressions give the same bits as
  //      on host storage, their results stay in HBM, and a chain of them
  //      leaves every tensor on the device
  {
    auto same = [](auto &h, auto &d) {  // d is brought back to the host
      if (!d.on_device()) return false;
      d.to_host();
      return h.indices() == d.indices() && h.data() == d.data() && h.linIdxSequence() == d.linIdxSequence() &&
             h.is_sorted() == d.is_sorted();
    };
    lnt::SparseMIA<double, 4> a(6, 5, 4, 3), b(4, 3, 6, 5), ad(6, 5, 4, 3), bd(4, 3, 6, 5);
    fill_random(a, 150, 51);
    fill_random(b, 200, 52);
    ad = a;
    bd = b;
    ad.to_device();
    bd.to_device();
    CHECK(ad.on_device() && ad.size() == a.size(), "to_device");
    bool thrown = false;
    try {
      (void)ad.indices();
    } catch (const lnt::MIAException &) {
      thrown = true;
    }
    CHECK(thrown, "host accessor of a device tensor did not throw");
    lnt::SparseMIA<double, 4> c(6, 5, 4, 3), cd(6, 5, 4, 3), e(6, 5, 4, 3), ed(6, 5, 4, 3);
    c(i, j, k, l) = a(i, j, k, l) + b(k, l, i, j);  // neither sorted: sortMerge
    cd(i, j, k, l) = ad(i, j, k, l) + bd(k, l, i, j);
    CHECK(cd.on_device() && ad.on_device() && bd.on_device(), "device operands or result left HBM");
    e(i, j, k, l) = c(i, j, k, l) - a(i, j, k, l);  // c sorted: scan merge
    ed(i, j, k, l) = cd(i, j, k, l) - ad(i, j, k, l);
    CHECK(same(e, ed), "device merge differs");
    CHECK(same(c, cd), "device sortMerge differs");
    std::array<size_t, 4> seq{{2, 0, 3, 1}};
    lnt::SparseMIA<double, 4> s = a, sd = a;
    sd.to_device();
    s.sort(seq);
    sd.sort(seq);
    CHECK(same(s, sd), "device sort(seq) differs");
    lnt::SparseMIA<double, 4> m1(6, 5, 4, 3), m1d(6, 5, 4, 3);
    m1(i, j, k, l) = a(!i, !j, !k, !l) * b(!k, !l, !i, !j);  // Hadamard
    m1d(i, j, k, l) = ad(!i, !j, !k, !l) * bd(!k, !l, !i, !j);
    CHECK(same(m1, m1d), "device Hadamard differs");
    lnt::SparseMIA<double, 3> x(6, 5, 4), y(4, 7, 3), xd(6, 5, 4), yd(4, 7, 3);
    lnt::SparseMIA<double, 4> z(6, 5, 7, 3), zd(6, 5, 7, 3);
    fill_random(x, 80, 53);
    fill_random(y, 60, 54);
    xd = x;
    yd = y;
    xd.to_device();
    z(i, j, m, n) = x(i, j, k) * y(k, m, n);  // contraction
    zd(i, j, m, n) = xd(i, j, k) * yd(k, m, n);  // y still on the host: moved to HBM
    CHECK(yd.on_device(), "host operand of a device product not moved to HBM");
    CHECK(same(z, zd), "device contraction differs");
    lnt::SparseMIA<double, 3> g(5, 6, 5), gd(5, 6, 5);
    fill_random(g, 90, 56);
    gd = g;
    gd.to_device();
    lnt::SparseMIA<double, 1> tr(6), trd(6);
    tr(j) = g(i, j, i);  // unary contraction
    trd(j) = gd(i, j, i);
    CHECK(same(tr, trd), "device unary contraction differs");
    lnt::DenseMIA<double, 2> wdense(4, 2);
    for (int64_t q = 0; q < 8; ++q) wdense.data_begin()[q] = 0.25 * (double)(q + 1);
    lnt::SparseMIA<double, 4> r(6, 5, 2, 4), rd(6, 5, 2, 4);
    r(i, j, m, k) = x(i, j, !k) * wdense(!k, m);  // sparse x dense
    rd(i, j, m, k) = xd(i, j, !k) * wdense(!k, m);
    CHECK(same(r, rd), "device sparse x dense differs");
    lnt::SparseMIA<double, 3> w(6, 5, 4), wd(6, 5, 4);
    fill_random(w, 60, 55);
    lnt::SparseMIA<double, 3> w0 = w;
    wd = w;
    wd.to_device();
    w(i, j, k) += x(i, j, k);
    wd(i, j, k) += xd(i, j, k);
    CHECK(same(w, wd), "device += differs");
    lnt::SparseMIA<double, 3> q1 = w0, q1d = w0;
    q1d.to_device();
    CHECK(q1 == w0 && q1d == w0, "device comparison");
    double h = x(i, j, k) * x(i, j, k), hd = xd(i, j, k) * xd(i, j, k);
    CHECK(h == hd, "device complete inner product differs");
    lnt::SparseMIA<double, 3> back = xd;
    back.to_host();
    CHECK(back.indices() == x.indices() && back.data() == x.data(), "to_device/to_host round trip");
  }
  if (failures) return 1;
  std::printf("C++ front end: all checks passed\n");
  return 0;
}
