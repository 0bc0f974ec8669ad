"""Full-size parity of the sort/merge-at-scale workload (the benchmark's own
inputs): two order-4 tensors over 256^4, 50M nonzeros each, 25M keys shared.
Every array the benchmark step produces is compared with the CPU oracle bit
for bit: the initial sorts (SparseMIABase::sort, SparseMIABase.h:701-734), the
2-pass change_sort_order to [0, 2, 1, 3] and back (SparseMIABase.h:590-645),
and A+B, A-B, A∘B (FunctionUtil.h:373-419, SparseMIABase.h:1717-1826)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host(t):
    return t.cpu().numpy()


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def test_sort_merge_at_scale_matches_oracle(ctx, orc):
    import torch

    import bench

    ak, av, bk, bv = bench.make_c2(ctx, torch)
    ns, no = bench.N_SHARED, bench.N_ONLY
    assert ak.shape[0] == bk.shape[0] == ns + no
    # Initial sorts: the generator's keys sorted by the oracle's MSD radix.
    keys = orc.gen_distinct_keys(ns + 2 * no, 1 << 32, bench.SEED_KEYS)
    ea_k, ea_v = orc.radix_sort_pairs(keys[: ns + no], orc.gen_uniform_vals(ns + no, bench.SEED_A), 1 << 32)
    eb_k, eb_v = orc.radix_sort_pairs(np.concatenate([keys[:ns], keys[ns + no:]]),
                                      orc.gen_uniform_vals(ns + no, bench.SEED_B), 1 << 32)
    del keys
    assert np.array_equal(host(ak), ea_k) and bits_equal(host(av), ea_v)
    assert np.array_equal(host(bk), eb_k) and bits_equal(host(bv), eb_v)
    perm = bench._perm()
    assert perm == [0, 2, 1, 3]
    ctx.set_profiling(True)
    for cur, new in ((bench.IDENT, perm), (perm, bench.IDENT)):
        ctx.change_sort_order(ak, av, bench.DIMS, cur, new, True)
        ctx.change_sort_order(bk, bv, bench.DIMS, cur, new, True)
        ea_k, ea_v = orc.sort_seq(ea_k, ea_v, bench.DIMS, cur, new, True)
        eb_k, eb_v = orc.sort_seq(eb_k, eb_v, bench.DIMS, cur, new, True)
        assert np.array_equal(host(ak), ea_k) and bits_equal(host(av), ea_v), (cur, new)
        assert np.array_equal(host(bk), eb_k) and bits_equal(host(bv), eb_v), (cur, new)
        if cur == bench.IDENT:
            for op in (0, 1, 2):
                ck, cv = ctx.merge(ak, av, bk, bv, op)
                xk, xv = orc.merge(ea_k, ea_v, eb_k, eb_v, op)
                assert len(xk) == (ns if op == 2 else ns + 2 * no)
                assert np.array_equal(host(ck), xk) and bits_equal(host(cv), xv), op
                del ck, cv, xk, xv
    # Two passes per sort: the low 16 bits (modes 0 and 2 keep their order).
    assert ctx.profile("onesweep")[0] == 8
    ctx.set_profiling(False)
