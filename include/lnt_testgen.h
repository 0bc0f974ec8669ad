/*
 * lnt_testgen.h -- synthetic-input generators that run on the GPU.
 *
 * No reference counterpart: the reference draws its test data from
 * boost::mt19937 seeded by time(0) (multi_index_array/LibMIAUtil.h:43-46).
 * These counter-based generators are bit-identical to the CPU oracle's
 * (oracle/lnt_oracle.c lnto_permute_index / lnto_uniform_pm1), so a benchmark
 * can build inputs of hundreds of millions of entries directly in HBM and a
 * test can rebuild the same inputs on the CPU.
 */
#ifndef LNT_TESTGEN_H
#define LNT_TESTGEN_H

#include "lnt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* keys[t] = permute_index(offset + t, domain, seed): distinct values in
 * [0, domain) for distinct t < domain.  Device pointer. */
int lnt_testgen_distinct_keys(lnt_context ctx, int64_t *keys, int64_t n,
                              uint64_t domain, uint64_t seed, uint64_t offset);
/* vals[t] = uniform_pm1(seed, offset + t) in [-1, 1).  Device pointer. */
int lnt_testgen_uniform(lnt_context ctx, double *vals, int64_t n, uint64_t seed,
                        uint64_t offset);

/* A batch of instances: instance i (0 <= i < count) owns entries
 * [offsets[i], offsets[i+1]) (host array, offsets[0] = 0) and gets
 * keys permute_index(t, domains[i], key_seed + i) and values
 * uniform_pm1(val_seed + i, t), t = 0, 1, ... (each instance as one
 * lnt_testgen_distinct_keys / lnt_testgen_uniform call with offset 0).
 * keys or vals may be NULL.  Device pointers; synchronises. */
int lnt_testgen_batch(lnt_context ctx, int64_t *keys, double *vals, const int64_t *offsets,
                      const uint64_t *domains, int64_t count, uint64_t key_seed, uint64_t val_seed);

#ifdef __cplusplus
}
#endif
#endif
