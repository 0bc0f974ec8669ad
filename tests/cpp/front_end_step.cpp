// The sort/merge-at-scale step written against the C++ front end, the way a
// reference user writes it (A.sort(seq); C(i,j,k,l) = A(i,j,k,l) + B(i,j,k,l);
// Hadamard), timed with the tensors resident in HBM ("device") or in the
// reference's host vectors ("host", every call staged over PCIe).  Inputs:
// order-4 tensors over 256^4, 25M shared + 25M own keys each, from the same
// counter-based generators and seeds as bench.py.  Prints one JSON line.
//   front_end_step device|host STEPS WARMUP PERM0 PERM1 PERM2 PERM3
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "lnt/SparseMIA.hpp"
#include "lnt_testgen.h"

using Clock = std::chrono::steady_clock;

static void make(lnt::SparseMIA<double, 4> &m, int64_t n_shared, int64_t n_own, uint64_t own_offset,
                 uint64_t val_seed) {
  lnt_context ctx = lnt::default_context().get();
  const int64_t n = n_shared + n_own;
  void *k = nullptr, *v = nullptr;
  lnt::check(lnt_alloc(ctx, n * 8, LNT_MEM_DEVICE, &k));
  lnt::check(lnt_alloc(ctx, n * 8, LNT_MEM_DEVICE, &v));
  int64_t *keys = (int64_t *)k;
  lnt::check(lnt_testgen_distinct_keys(ctx, keys, n_shared, uint64_t(1) << 32, 2024, 0));
  lnt::check(lnt_testgen_distinct_keys(ctx, keys + n_shared, n_own, uint64_t(1) << 32, 2024, own_offset));
  lnt::check(lnt_testgen_uniform(ctx, (double *)v, n, val_seed, 0));
  m.assign_device(keys, (double *)v, n, std::array<int, 4>{{0, 1, 2, 3}}, false);
  m.sort();
}

int main(int argc, char **argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: %s device|host STEPS WARMUP P0 P1 P2 P3\n", argv[0]);
    return 2;
  }
  const bool device = std::strcmp(argv[1], "device") == 0;
  const int steps = std::atoi(argv[2]), warmup = std::atoi(argv[3]);
  const std::array<int, 4> ident{{0, 1, 2, 3}}, perm{{std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]),
                                                      std::atoi(argv[7])}};
  NTINDEX i; NTINDEX j; NTINDEX k; NTINDEX l;
  const int64_t ns = 25000000, no = 25000000;
  lnt::SparseMIA<double, 4> a(256, 256, 256, 256), b(256, 256, 256, 256);
  make(a, ns + no, 0, 0, 11);
  make(b, ns, no, ns + no, 12);
  if (!device) {
    a.to_host();
    b.to_host();
  }
  lnt::SparseMIA<double, 4> c(256, 256, 256, 256), d(256, 256, 256, 256);
  size_t nc = 0, nd = 0;
  auto step = [&](int s) {
    const auto &seq = (s % 2 == 0) ? perm : ident;
    a.sort(seq);
    b.sort(seq);
    c(i, j, k, l) = a(i, j, k, l) + b(i, j, k, l);
    d(i, j, k, l) = a(!i, !j, !k, !l) * b(!i, !j, !k, !l);
    nc = c.size();
    nd = d.size();
  };
  for (int s = 0; s < warmup; ++s) step(s);
  lnt::check(lnt_context_synchronize(lnt::default_context().get()));
  const auto t0 = Clock::now();
  for (int s = 0; s < steps; ++s) step(warmup + s);
  lnt::check(lnt_context_synchronize(lnt::default_context().get()));
  const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count() / steps;
  std::printf("{\"storage\": \"%s\", \"steps\": %d, \"ms_per_step\": %.4f, \"keys_per_s\": %.6g, "
              "\"union_nnz\": %zu, \"hadamard_nnz\": %zu, \"on_device\": %s}\n",
              device ? "device" : "host", steps, ms, (double)(2 * (ns + no)) / (ms * 1e-3), nc, nd,
              c.on_device() ? "true" : "false");
  return (nc == (size_t)(ns + 2 * no) && nd == (size_t)ns) ? 0 : 1;
}
