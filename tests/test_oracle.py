"""Pins the CPU oracle (oracle/lnt_oracle.c) against known answers, numpy
library routines, dense brute force and invariants -- never against a retyped
copy of its own formulas.  CPU only."""
import itertools

import numpy as np
import pytest


# ------------------------------------------------------------- generators
def test_splitmix64_known_answers(orc):
    # Published first outputs of SplitMix64 with state 0 (Vigna's splitmix64.c).
    assert orc.rand_u64(0, 0) == 0xE220A8397B1DCDAF
    assert orc.rand_u64(0, 1) == 0x6E789E6AA1B965F4
    assert orc.rand_u64(0, 2) == 0x06C45D188009454F


def test_uniform_is_exact_53_bit_grid(orc):
    for i in range(2000):
        v = orc.uniform_pm1(12345, i)
        assert -1.0 <= v < 1.0
        m = (v + 1.0) * 2.0 ** 52
        assert m == int(m) and int(m) == orc.rand_u64(12345, i) >> 11


@pytest.mark.parametrize("bits", [2, 4, 6, 8, 10, 12, 14, 16])
def test_feistel_is_a_bijection(orc, bits):
    import ctypes

    out = [orc.feistel(x, bits, 99) for x in range(1 << bits)]
    assert sorted(out) == list(range(1 << bits))
    assert out != list(range(1 << bits))


@pytest.mark.parametrize("domain", [1, 2, 3, 5, 1000, 4097, 12345])
def test_permute_index_is_a_bijection(orc, domain):
    k = orc.gen_distinct_keys(domain, domain, 7)
    assert sorted(k.tolist()) == list(range(domain))


def test_distinct_keys_prefix(orc):
    k = orc.gen_distinct_keys(10486, 1 << 20, 1)
    assert len(np.unique(k)) == len(k) and k.min() >= 0 and k.max() < (1 << 20)
    k2 = orc.gen_distinct_keys(100, 1 << 20, 1, offset=50)
    assert np.array_equal(k[50:150], k2)


def test_random_perm(orc):
    for seed in range(50):
        p = orc.random_perm(6, seed)
        assert sorted(p.tolist()) == list(range(6))


def test_gen_coords_distinct_and_skewed(orc):
    dims = [64, 64, 32]
    k, v = orc.gen_coords(5000, dims, [1, 0, 0], zipf_s=1.1, value_kind=1, seed=3)
    assert len(np.unique(k)) == 5000 and k.min() >= 0 and k.max() < 64 * 64 * 32
    i = k % 64
    # Zipf(1.1) over 64 values puts far more mass on 0 than uniform (1/64).
    assert (i == 0).mean() > 5.0 / 64
    j = (k // 64) % 64
    assert abs((j == 0).mean() - 1.0 / 64) < 0.02
    assert abs(v.mean()) < 0.1 and 0.8 < v.std() < 1.2
    k2, v2 = orc.gen_coords(5000, dims, [1, 0, 0], zipf_s=1.1, value_kind=1, seed=3)
    assert np.array_equal(k, k2) and np.array_equal(v, v2)


# ------------------------------------------------------------- index math
def _np_relin(keys, dims, cur, new):
    dims = np.asarray(dims)
    subs_r = np.unravel_index(keys, [int(dims[m]) for m in cur], order="F")
    sub = [None] * len(dims)
    for t, m in enumerate(cur):
        sub[m] = subs_r[t]
    return np.ravel_multi_index([sub[m] for m in new], [int(dims[m]) for m in new], order="F")


def test_change_linidx_exhaustive_small(orc):
    dims = [2, 3, 4, 5]
    keys = np.arange(120, dtype=np.int64)
    perms = list(itertools.permutations(range(4)))
    for cur in perms[::3]:
        for new in perms:
            got = orc.change_linidx(keys, dims, cur, new)
            assert np.array_equal(got, _np_relin(keys, dims, cur, new)), (cur, new)


def test_change_linidx_round_trip_large(orc):
    rng = np.random.default_rng(0)
    dims = [4096, 4096, 512, 64, 3]  # 2^42 * 3 cells
    total = int(np.prod(dims, dtype=object))
    keys = np.array([int(x) for x in rng.integers(0, total, 5000, dtype=np.int64)], dtype=np.int64)
    cur, new = [0, 1, 2, 3, 4], [3, 0, 4, 2, 1]
    fwd = orc.change_linidx(keys, dims, cur, new)
    # Python big-int closed form.
    for key, f in zip(keys[:200].tolist(), fwd[:200].tolist()):
        sub = [0] * 5
        k = key
        for m in cur:
            sub[m] = k % dims[m]
            k //= dims[m]
        e, mult = 0, 1
        for m in new:
            e += sub[m] * mult
            mult *= dims[m]
        assert e == f
    back = orc.change_linidx(fwd, dims, new, cur)
    assert np.array_equal(back, keys)


def test_ind2sub_sub2ind(orc):
    dims = [3, 5, 7]
    seq = [2, 0, 1]
    for key in range(105):
        sub = orc.ind2sub(key, dims, seq)
        assert orc.sub2ind(sub, dims, seq) == key
        assert np.ravel_multi_index([sub[m] for m in seq], [dims[m] for m in seq], order="F") == key


# ---------------------------------------------------------------- sorting
@pytest.mark.parametrize("n", [0, 1, 2, 49, 50, 51, 2999, 3000, 3001, 20000, 200000])
@pytest.mark.parametrize("bound", [1 << 20, (1 << 33) + 17, 1 << 62])
def test_radix_sort_matches_argsort(orc, n, bound):
    n = min(n, bound)
    keys = orc.gen_distinct_keys(n, bound, n + 11)
    vals = orc.gen_uniform_vals(n, 5)
    k, v = orc.radix_sort_pairs(keys, vals, bound)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v.view(np.int64), vals[order].view(np.int64))


def test_radix_sort_with_duplicates(orc):
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 700, 50000).astype(np.int64)
    vals = rng.standard_normal(50000)
    k, v = orc.radix_sort_pairs(keys, vals, 700)
    assert np.all(np.diff(k) >= 0)
    # (key, value) multiset preserved
    a = sorted(zip(keys.tolist(), vals.tolist()))
    b = sorted(zip(k.tolist(), v.tolist()))
    assert a == b


def test_introsort(orc):
    keys = orc.gen_distinct_keys(5000, 1 << 30, 4)
    vals = orc.gen_uniform_vals(5000, 4)
    k, v = orc.introsort_pairs(keys, vals)
    o = np.argsort(keys)
    assert np.array_equal(k, keys[o]) and np.array_equal(v, vals[o])


@pytest.mark.parametrize("n", [100, 2999, 3000, 20000])
@pytest.mark.parametrize("is_sorted", [False, True])
def test_sort_seq_is_relinearise_then_sort(orc, n, is_sorted):
    """sparse_mia_functions.cpp:100-202: sort(seq) == change_linIdx_sequence(seq)+sort()."""
    dims = [32, 32, 32, 32]
    keys = orc.gen_distinct_keys(n, 1 << 20, 9)
    vals = orc.gen_uniform_vals(n, 9)
    cur = [0, 1, 2, 3]
    if is_sorted:
        keys, vals = orc.radix_sort_pairs(keys, vals, 1 << 20)
    for new in ([2, 0, 3, 1], [3, 2, 1, 0], [0, 1, 2, 3], [1, 0, 2, 3]):
        k, v = orc.sort_seq(keys, vals, dims, cur, new, is_sorted)
        rk = _np_relin(keys, dims, cur, new)
        o = np.argsort(rk)
        assert np.array_equal(k, rk[o]) and np.array_equal(v, vals[o])


@pytest.mark.parametrize("seed", range(6))
def test_radix_shuffle_sorts_power_of_two_shapes(orc, seed):
    """The restated RadixShuffle (the reference's sort(seq) on sorted input,
    SparseMIABase.h:622-634) gives numpy's argsort of the relinearised keys on
    power-of-two dims, where its divisor approximation is exact: random orders
    2-6 (sizes 2 to 128), random sequences, 3,000 to 300,000 keys (every branch of its cost
    model: find-first, sort-first, shuffle and straight runs), plus the
    sort/merge-at-scale permutation both ways."""
    rng = np.random.default_rng(100 + seed)
    cases = []
    while len(cases) < 12:
        order = int(rng.integers(2, 7))
        dims = [int(2 ** rng.integers(1, 8)) for _ in range(order)]
        tot = int(np.prod(dims))
        if tot < 4000:
            continue
        n = int(rng.integers(3000, min(tot, 300000) + 1))
        cases.append((dims, [int(x) for x in rng.permutation(order)], [int(x) for x in rng.permutation(order)], n))
    if seed == 0:
        cases += [([256] * 4, [0, 1, 2, 3], [0, 2, 1, 3], 400000), ([256] * 4, [0, 2, 1, 3], [0, 1, 2, 3], 400000)]
    for i, (dims, cur, new, n) in enumerate(cases):
        keys = orc.gen_distinct_keys(n, int(np.prod(dims)), 7 * seed + i)
        vals = orc.gen_uniform_vals(n, i)
        keys, vals = orc.radix_sort_pairs(_np_relin(keys, dims, list(range(len(dims))), cur), vals,
                                          int(np.prod(dims)))
        k, v = orc.sort_seq_shuffle(keys, vals, dims, cur, new, True)
        rk = _np_relin(keys, dims, cur, new)
        o = np.argsort(rk)
        assert np.array_equal(k, rk[o]) and np.array_equal(v.view(np.int64), vals[o].view(np.int64)), (dims, cur, new)


def test_radix_shuffle_defect_on_other_dims(orc):
    """RadixShuffle plans a sort stage's passes from the stage's size but takes
    digits as key >> floor(log2(divisor)); with dims {19,25,35,3} and
    {2,3,0,1} -> {1,3,0,2} its one stage has divisor 25 and 1,995 values, so
    key >> 4 reaches 3,117 where one 11-bit pass keeps 2,048: bins mix and the
    reference leaves the keys unsorted.  The restatement follows the reference
    there (a version that sized its passes from the approximated range would
    sort them); the checker's sort_seq sorts fully."""
    dims, cur, new = [19, 25, 35, 3], [2, 3, 0, 1], [1, 3, 0, 2]
    keys = np.sort(orc.gen_distinct_keys(13495, int(np.prod(dims)), 5))
    vals = orc.gen_uniform_vals(13495, 12)
    k, _ = orc.sort_seq_shuffle(keys, vals, dims, cur, new, True)
    rk = np.sort(_np_relin(keys, dims, cur, new))
    assert np.array_equal(np.sort(k), rk)  # a permutation of the right keys ...
    assert not np.all(np.diff(k) > 0)      # ... not sorted
    k2, _ = orc.sort_seq(keys, vals, dims, cur, new, True)
    assert np.array_equal(k2, rk)


def test_radix_shuffle_defect_on_unit_dims(orc):
    """setupPermute (IndexUtil.h:675-728) drops a stage of size 1; when the
    next index is of the previous stage's kind, the stage is reused but the
    index is not folded into it (its dims do not multiply the stage size nor
    lower its divisor), so the stage plan misses it.  dims {4,4,1,32,32,8},
    {0,4,3,5,1,2} -> {3,4,2,5,0,1}: new position 1 (size 32) is left out and
    the reference's output is unsorted; the restatement follows it."""
    dims, cur, new = [4, 4, 1, 32, 32, 8], [0, 4, 3, 5, 1, 2], [3, 4, 2, 5, 0, 1]
    tot = int(np.prod(dims))
    keys = np.sort(_np_relin(orc.gen_distinct_keys(47635, tot, 14), dims, list(range(6)), cur))
    vals = orc.gen_uniform_vals(47635, 0)
    k, _ = orc.sort_seq_shuffle(keys, vals, dims, cur, new, True)
    rk = np.sort(_np_relin(keys, dims, cur, new))
    assert np.array_equal(np.sort(k), rk) and not np.all(np.diff(k) > 0)
    k2, _ = orc.sort_seq(keys, vals, dims, cur, new, True)
    assert np.array_equal(k2, rk)


def _v2(x):
    b = 0
    while x % 2 == 0 and b < 63:
        x //= 2
        b += 1
    return b


@pytest.mark.parametrize("dims,cur,new", [
    ([256, 256, 256, 256], [0, 1, 2, 3], [0, 2, 1, 3]),
    ([12, 5, 7, 9], [0, 1, 2, 3], [0, 2, 1, 3]),
    ([6, 4, 10], [2, 0, 1], [2, 1, 0]),
    ([4, 4, 4, 4, 4], [1, 0, 2, 3, 4], [1, 0, 4, 3, 2]),
    ([24, 6, 5], [0, 1, 2], [0, 2, 1]),
    ([8, 8, 8, 8], [0, 1, 2, 3], [1, 3, 0, 2]),
])
def test_sorted_low_modes_lemma(orc, dims, cur, new):
    """Keys sorted in `cur`, relinearised into `new`: when the first J modes of
    `new` appear in `cur` in the same relative order and their dims multiply
    to P = 2^b * odd, a stable sort on (key >> b) alone already gives the full
    order (the low b bits need no radix pass).  Plausible mistakes break it on
    the same data: skipping P's bit length when P is not a power of two, and
    counting the first mode out of order as in order."""
    total = int(np.prod(dims))
    keys = np.sort(orc.gen_distinct_keys(min(20000, total // 2), total, 3))
    rk = orc.change_linidx(keys, dims, cur, new)
    pos = {m: t for t, m in enumerate(cur)}
    p, j = 1, 0
    while j < len(dims) and (j == 0 or pos[new[j]] > pos[new[j - 1]]):
        p *= dims[new[j]]
        j += 1
    b = _v2(p)
    full = np.sort(rk)
    assert np.array_equal(rk[np.argsort(rk >> b, kind="stable")], full)
    if p & (p - 1):
        wrong = int(p - 1).bit_length()
        assert not np.array_equal(rk[np.argsort(rk >> wrong, kind="stable")], full)
    wrong = _v2(p * dims[new[j]])
    if wrong > b:
        assert not np.array_equal(rk[np.argsort(rk >> wrong, kind="stable")], full)


def test_compare_sparse_known_answers(orc):
    """compare_with_sparse (SparseMIABase.h:1243-1296) on hand-built cases whose
    answers follow from its definition: entries with |v| <= 1.11e-16 are absent,
    the rest must pair up key for key with the predicate true."""
    k = np.array([2, 5, 9, 11], dtype=np.int64)
    v = np.array([1.5, -0.25, 3.0, 0.75])
    c = orc.compare_sparse
    assert c(k, v, k, v)
    assert not c(k, v, k, v + np.array([0, 0, 1e-9, 0]))                    # exact
    assert c(k, v, k, v + np.array([0, 0, 1e-9, 0]), 1, 1e-6)                # relative, |a|>=1
    assert not c(k, v, k, v + np.array([0, 2e-6, 0, 0]), 1, 1e-6)            # absolute, |a|<1
    assert c(k, v, k, v + np.array([0, 5e-7, 0, 0]), 1, 1e-6)
    assert c(k, v, np.array([2, 5, 7, 9, 11]), np.array([1.5, -0.25, 1e-17, 3.0, 0.75]))   # absent extra
    assert c(k, v, np.array([2, 5, 7, 9, 11]), np.array([1.5, -0.25, -1.11e-16, 3.0, 0.75]))
    assert not c(k, v, np.array([2, 5, 7, 9, 11]), np.array([1.5, -0.25, 2e-16, 3.0, 0.75]))
    assert not c(k, v, k[:3], v[:3])                                         # missing key
    assert c(np.append(k, 20), np.append(v, 0.0), k, v)                      # trailing zero
    assert not c(k, v, np.array([2, 5, 10, 11]), v)                          # key differs
    assert not c(k, np.array([1.5, np.nan, 3.0, 0.75]), k, np.array([1.5, np.nan, 3.0, 0.75]))
    assert c(np.zeros(0, np.int64), np.zeros(0), np.zeros(0, np.int64), np.zeros(0))


def test_mex_merge_known_answers(orc):
    """SparseMergeMex (src/MATLAB/mex/sparse_merge.cpp:16-85): with both
    operands present it is the general merge; an empty operand takes the mex's
    shortcuts, told apart from the general merge by signed zeros: a subtraction
    from nothing is -1*b (so +0.0 -> -0.0), where the merge computes 0 - b
    (+0.0 -> +0.0), and an empty B hands A back unchanged (-0.0 stays)."""
    sb = lambda a: np.signbit(np.asarray(a)).tolist()
    k = np.array([1, 4, 6], dtype=np.int64)
    v = np.array([0.0, -0.0, 2.0])
    e_k, e_v = np.zeros(0, np.int64), np.zeros(0)
    ck, cv = orc.mex_merge(e_k, e_v, k, v, 1)
    assert ck.tolist() == [1, 4, 6] and cv.tolist() == [0.0, 0.0, -2.0] and sb(cv) == [True, False, True]
    gk, gv = orc.merge(e_k, e_v, k, v, 1)  # the general merge: 0 - b
    assert gk.tolist() == [1, 4, 6] and sb(gv) == [False, False, True]
    ck, cv = orc.mex_merge(e_k, e_v, k, v, 0)
    assert sb(cv) == [False, True, False]
    gk, gv = orc.merge(e_k, e_v, k, v, 0)  # 0 + b loses -0.0
    assert sb(gv) == [False, False, False]
    ck, cv = orc.mex_merge(k, v, e_k, e_v, 1)
    assert ck.tolist() == [1, 4, 6] and sb(cv) == [False, True, False]
    assert len(orc.mex_merge(e_k, e_v, e_k, e_v, 0)[0]) == 0
    # both present: the general merge, by hand
    ck, cv = orc.mex_merge(k, np.array([1.0, 2.0, 3.0]), np.array([0, 4, 9], np.int64),
                           np.array([5.0, 7.0, 1.0]), 1)
    assert ck.tolist() == [0, 1, 4, 6, 9] and cv.tolist() == [-5.0, 1.0, -5.0, 3.0, -1.0]


def test_collect_duplicates(orc):
    keys = np.array([1, 1, 2, 5, 5, 5, 9], dtype=np.int64)
    vals = np.array([1.0, 2.0, 3.0, 1.0, 1.0, 1.0, -4.0])
    k, v = orc.collect_duplicates(keys, vals)
    assert k.tolist() == [1, 2, 5, 9] and v.tolist() == [3.0, 3.0, 3.0, -4.0]


def test_sort_merge_zero_rule_known_answers(orc):
    """SparseMIA::sortMerge + collect_duplicates (SparseMIA.h:628-665,
    833-887) by hand.  The loop keeps the first entry of the sorted array
    unconditionally; any later key opens only on a non-zero entry (a zero
    first entry is skipped and the key's next entry is tried as new); equal
    keys add.  Subtraction copies -1*b and then adds."""
    # A={1:0, 3:1, 5:2, 7:0}, B={1:2, 4:0, 5:-2, 7:-0}: 1 is the smallest key
    # (kept: 0+2); 4 holds only a zero and 7 only zeros (both vanish); 5 sums
    # to an explicit zero, which stays (it opened on a non-zero entry).
    ck, cv = orc.sort_merge(np.array([3, 1, 7, 5]), np.array([1.0, 0.0, 0.0, 2.0]),
                            np.array([5, 1, 7, 4]), np.array([-2.0, 2.0, -0.0, 0.0]), orc.OP_ADD, 8)
    assert ck.tolist() == [1, 3, 5] and cv.tolist() == [2.0, 1.0, 0.0]
    # The smallest key stays even when every entry of it is zero, and keeps
    # the sum's sign: 0.0 + 0.0 = +0.0 whichever entry comes first.
    ck, cv = orc.sort_merge(np.array([0, 6]), np.array([0.0, 3.0]), np.array([0]), np.array([0.0]), orc.OP_ADD, 8)
    assert ck.tolist() == [0, 6] and cv.tolist() == [0.0, 3.0]
    assert not np.signbit(cv[0])
    # -0.0 + -0.0 = -0.0 on the smallest key (both zero, kept).
    ck, cv = orc.sort_merge(np.array([2]), np.array([-0.0]), np.array([2]), np.array([-0.0]), orc.OP_ADD, 8)
    assert ck.tolist() == [2] and np.signbit(cv[0])
    # Subtraction negates B first: 3 - 3 = 0 kept, B-only 4 -> -5; a B-only
    # key whose value is 0 becomes -1*0 = -0.0, a zero, and vanishes.
    ck, cv = orc.sort_merge(np.array([2, 3]), np.array([1.0, 3.0]), np.array([4, 3, 6]), np.array([5.0, 3.0, 0.0]),
                            orc.OP_SUB, 8)
    assert ck.tolist() == [2, 3, 4] and cv.tolist() == [1.0, 0.0, -5.0]
    # A zero A entry followed by a non-zero B entry of the same (not smallest)
    # key opens the key with the B value alone: 0 is skipped, not added.
    ck, cv = orc.sort_merge(np.array([1, 9]), np.array([1.0, 0.0]), np.array([9]), np.array([-0.0]), orc.OP_SUB, 16)
    assert ck.tolist() == [1] and cv.tolist() == [1.0]
    ck, cv = orc.sort_merge(np.array([1, 9]), np.array([1.0, 0.0]), np.array([9]), np.array([-7.5]), orc.OP_SUB, 16)
    assert ck.tolist() == [1, 9] and cv.tolist() == [1.0, 7.5]


# ---------------------------------------------------------------- merging
def _dense(keys, vals, size):
    d = np.zeros(size)
    d[keys] = vals
    return d


def _mk_pair(orc, na, nb, overlap, domain, seed):
    """Sorted A, B with `overlap` shared keys (distinct keys from one bijection)."""
    keys = orc.gen_distinct_keys(na + nb - overlap, domain, seed)
    ak = np.sort(keys[:na])
    bk = np.sort(np.concatenate([keys[:overlap], keys[na:]]))
    av = orc.gen_uniform_vals(na, seed + 1)
    bv = orc.gen_uniform_vals(len(bk), seed + 2)
    return ak, av, bk, bv


@pytest.mark.parametrize("na,nb,ov", [(0, 0, 0), (0, 50, 0), (50, 0, 0), (300, 200, 100), (1000, 1000, 1000), (5000, 3000, 0)])
def test_merge_union_and_hadamard_dense(orc, na, nb, ov):
    domain = 1 << 16
    ak, av, bk, bv = _mk_pair(orc, na, nb, ov, domain, 21)
    A, B = _dense(ak, av, domain), _dense(bk, bv, domain)
    supp_union = np.union1d(ak, bk)
    for op, fn in ((orc.OP_ADD, np.add), (orc.OP_SUB, np.subtract)):
        ck, cv = orc.merge(ak, av, bk, bv, op)
        assert np.array_equal(ck, supp_union)
        assert np.array_equal(cv, fn(A, B)[ck])
    ck, cv = orc.merge(ak, av, bk, bv, orc.OP_HADAMARD)
    assert np.array_equal(ck, np.intersect1d(ak, bk))
    assert np.array_equal(cv, (A * B)[ck])


def test_merge_signed_zero_rules(orc):
    # B-only entries are op(0, b) (FunctionUtil.h:393,405): 0 + -0 = +0, 0 - +0 = +0,
    # while A-only entries are copied unchanged.
    ak = np.array([1], dtype=np.int64)
    av = np.array([-0.0])
    bk = np.array([2, 3], dtype=np.int64)
    bv = np.array([-0.0, 0.0])
    ck, cv = orc.merge(ak, av, bk, bv, orc.OP_ADD)
    assert ck.tolist() == [1, 2, 3]
    assert np.signbit(cv).tolist() == [True, False, False]
    ck, cv = orc.merge(ak, av, bk, bv, orc.OP_SUB)
    # 0 - (+0) is +0, where a negation (-b) would give -0.
    assert np.signbit(cv).tolist() == [True, False, False]


def test_outer_inter_product_dense(orc):
    rng = np.random.default_rng(2)
    L, R, T = 6, 5, 7
    A = np.where(rng.random((L, T)) < 0.4, rng.standard_normal((L, T)), 0.0)
    B = np.where(rng.random((R, T)) < 0.4, rng.standard_normal((R, T)), 0.0)
    ak = np.flatnonzero(A.T.ravel() != 0).astype(np.int64)  # key = l + L*t
    av = A.T.ravel()[ak]
    bk = np.flatnonzero(B.T.ravel() != 0).astype(np.int64)
    bv = B.T.ravel()[bk]
    ck, cv = orc.outer_inter_product(ak, av, L, bk, bv, R)
    C = np.einsum("it,jt->ijt", A, B)
    Cf = C.transpose(2, 1, 0).ravel()  # key = i + L*j + L*R*t
    assert np.array_equal(ck, np.flatnonzero(Cf != 0))
    assert np.array_equal(cv, Cf[ck])
    assert np.all(np.diff(ck) > 0)


# -------------------------------------------------------- lattice product
def _lattice_from_dense(D):
    h, w, d = D.shape
    flat = D.transpose(2, 1, 0).ravel()  # key = r + h*c + h*w*t
    k = np.flatnonzero(flat != 0).astype(np.int64)
    return k, flat[k]


@pytest.mark.parametrize("shape", [(20, 30, 1), (7, 40, 3), (1, 1, 1), (33, 1, 2), (40, 25, 4)])
@pytest.mark.parametrize("density", [0.05, 0.3])
def test_lattice_mult_dense(orc, shape, density):
    rng = np.random.default_rng(hash((shape, density)) & 0xFFFF)
    h, w, d = shape
    bw = 17
    A = np.where(rng.random((h, w, d)) < density, rng.uniform(-1, 1, (h, w, d)), 0.0)
    B = np.where(rng.random((w, bw, d)) < density, rng.uniform(-1, 1, (w, bw, d)), 0.0)
    ak, av = _lattice_from_dense(A)
    bk, bv = _lattice_from_dense(B)
    ck, cv = orc.lattice_mult(ak, av, (h, w, d), bk, bv, (w, bw, d))
    C = np.einsum("rkt,kct->rct", A, B)
    S = np.einsum("rkt,kct->rct", (A != 0).astype(np.int64), (B != 0).astype(np.int64))
    flatC = C.transpose(2, 1, 0).ravel()
    flatS = S.transpose(2, 1, 0).ravel()
    assert np.array_equal(ck, np.flatnonzero(flatS))  # structural nonzeros, sorted
    assert orc.allclose_fuzzy(cv, flatC[ck], 1e-12)
    nprod = int(S.sum())
    assert orc.lattice_nprod(ak, (h, w, d), bk, (w, bw, d)) == nprod


@pytest.mark.parametrize("method", ["spa", "mtt", "outer"])
@pytest.mark.parametrize("shape", [(20, 30, 1), (7, 40, 3), (1, 1, 1), (33, 1, 2), (40, 25, 4), (300, 6, 2)])
@pytest.mark.parametrize("density", [0.05, 0.3])
def test_lattice_mult_methods_dense(orc, method, shape, density):
    """csc_times, mtt_times and outer_times against dense einsum (values) and
    the boolean product (the structural index set, zero sums kept)."""
    rng = np.random.default_rng(hash((method, shape, density)) & 0xFFFF)
    h, w, d = shape
    bw = 13
    A = np.where(rng.random((h, w, d)) < density, rng.uniform(-1, 1, (h, w, d)), 0.0)
    B = np.where(rng.random((w, bw, d)) < density, rng.uniform(-1, 1, (w, bw, d)), 0.0)
    if h > 100:
        A[::3] = 0.0  # rows absent from A: the compressed row space of mtt_times is smaller than h
    ak, av = _lattice_from_dense(A)
    bk, bv = _lattice_from_dense(B)
    ck, cv = orc.lattice_mult_by(method, ak, av, (h, w, d), bk, bv, (w, bw, d))
    C = np.einsum("rkt,kct->rct", A, B)
    S = np.einsum("rkt,kct->rct", (A != 0).astype(np.int64), (B != 0).astype(np.int64))
    flatC = C.transpose(2, 1, 0).ravel()
    flatS = S.transpose(2, 1, 0).ravel()
    assert np.array_equal(ck, np.flatnonzero(flatS))
    assert orc.allclose_fuzzy(cv, flatC[ck], 1e-12)


def test_lattice_mult_spa_summation_order(orc):
    """The sparse accumulator adds a row's products in ascending B row order,
    the first one unadded: A(0,:) = [1, 1e16, -1e16] times B(:,0) = [1, 1, 1]
    gives (1 + 1e16) - 1e16 = 0 (1e16 + 1 rounds to 1e16), where descending
    order would give 1.  mtt_times performs the same additions; outer_times
    sums in the order its radix sort leaves, which on this input (one key
    three times) keeps the append order."""
    h, w, bw = 2, 3, 1
    A = np.zeros((h, w, 1))
    A[0, :, 0] = [1.0, 1e16, -1e16]
    A[1, 1, 0] = 2.0
    B = np.ones((w, bw, 1))
    ak, av = _lattice_from_dense(A)
    bk, bv = _lattice_from_dense(B)
    for method in ("spa", "mtt", "outer"):
        ck, cv = orc.lattice_mult_by(method, ak, av, (h, w, 1), bk, bv, (w, bw, 1))
        assert ck.tolist() == [0, 1] and cv.tolist() == [0.0, 2.0], method
    # csc_times and mtt_times: the same operations, the same bits.
    rng = np.random.default_rng(5)
    A = np.where(rng.random((60, 30, 2)) < 0.3, rng.standard_normal((60, 30, 2)) * 10.0 ** rng.integers(-8, 8, (60, 30, 2)), 0.0)
    B = np.where(rng.random((30, 20, 2)) < 0.3, rng.standard_normal((30, 20, 2)), 0.0)
    A[::4] = 0.0
    ak, av = _lattice_from_dense(A)
    bk, bv = _lattice_from_dense(B)
    k1, v1 = orc.lattice_mult_by("spa", ak, av, (60, 30, 2), bk, bv, (30, 20, 2))
    k2, v2 = orc.lattice_mult_by("mtt", ak, av, (60, 30, 2), bk, bv, (30, 20, 2))
    assert np.array_equal(k1, k2) and np.array_equal(v1.view(np.int64), v2.view(np.int64))


def _dense_mia(keys, vals, dims):
    D = np.zeros(int(np.prod(dims)))
    D[keys] = vals
    return D.reshape(dims, order="F")


@pytest.mark.parametrize(
    "la,lb,dims_a,dims_b",
    [
        ("ijkl", "klmn", [5, 4, 6, 3], [6, 3, 4, 5]),  # the c1 pattern
        ("ijk", "klm", [7, 6, 5], [5, 8, 3]),  # the c3 pattern
        ("kij", "jlk", [4, 5, 6], [6, 7, 4]),  # non-consecutive
        ("iJk", "kJm", [5, 3, 4], [4, 3, 6]),  # one inter index
        ("Ij", "jI", [6, 5], [5, 6]),  # inter + inner
        ("ij", "kl", [3, 4], [5, 2]),  # pure outer
        ("IJ", "IJ", [4, 5], [4, 5]),  # pure Hadamard
    ],
)
def test_contract_dense(orc, la, lb, dims_a, dims_b):
    rng = np.random.default_rng(len(la) * 7 + len(lb))
    na = max(1, int(0.3 * np.prod(dims_a)))
    nb = max(1, int(0.3 * np.prod(dims_b)))
    ak = orc.gen_distinct_keys(na, int(np.prod(dims_a)), 5)
    av = orc.gen_uniform_vals(na, 6)
    bk = orc.gen_distinct_keys(nb, int(np.prod(dims_b)), 7)
    bv = orc.gen_uniform_vals(nb, 8)
    A = _dense_mia(ak, av, dims_a)
    B = _dense_mia(bk, bv, dims_b)
    ck, cv, dims_c = orc.contract(ak, av, dims_a, list(range(len(dims_a))), False, la,
                                  bk, bv, dims_b, list(range(len(dims_b))), False, lb)
    # C's modes: [A outer, B outer, inter] (MIA_Expr.h:493-501)
    a_out = [c for c in la if c not in lb]
    b_out = [c for c in lb if c not in la]
    inter = [c for c in la if c.isupper()]
    out = "".join(a_out + b_out + inter).lower()
    C = np.einsum(la.lower() + "," + lb.lower() + "->" + out, A, B)
    S = np.einsum(la.lower() + "," + lb.lower() + "->" + out, (A != 0) * 1, (B != 0) * 1)
    assert list(dims_c) == list(C.shape)
    flatC, flatS = C.ravel(order="F"), S.ravel(order="F")
    assert np.array_equal(ck, np.flatnonzero(flatS))
    assert orc.allclose_fuzzy(cv, flatC[ck], 1e-12)


def test_contract_plan_rejects_bad_labels(orc):
    with pytest.raises(ValueError):
        orc.contract_plan([2, 3], "ii", [3, 2], "ij")  # repeated label
    with pytest.raises(ValueError):
        orc.contract_plan([2, 3], "iJ", [3, 2], "ik")  # inter label on one side only
    with pytest.raises(ValueError):
        orc.contract_plan([2, 3], "ij", [4, 2], "jk")  # mismatched contracted dims


def test_is_equal_fuzzy(orc):
    assert orc.is_equal_fuzzy(0.5, 0.5 + 1e-13, 1e-12)
    assert not orc.is_equal_fuzzy(0.5, 0.5 + 1e-11, 1e-12)
    assert orc.is_equal_fuzzy(1e6, 1e6 * (1 + 5e-13), 1e-12)  # relative above 1
    assert not orc.is_equal_fuzzy(1e6, 1e6 * (1 + 5e-12), 1e-12)


# ------------------------------------------------- unary contraction/attraction
from unary_cases import UNARY_CASES, unary_groups as _unary_groups  # noqa: E402


@pytest.mark.parametrize("labels,dims,spec", UNARY_CASES)
def test_contract_attract_against_dense(orc, labels, dims, spec):
    """The reference's unary patterns (tests/MIA/mia_unary_test.cpp:52-171)
    against dense numpy einsum, from an input in a shuffled mode order."""
    group, kind = _unary_groups(labels)
    dom = int(np.prod(dims))
    n = dom // 3
    k = orc.gen_distinct_keys(n, dom, 17)
    v = orc.gen_uniform_vals(n, 18) * 5
    seq = list(np.random.default_rng(len(labels)).permutation(len(dims)))
    dense = np.zeros(dom)
    dense[k] = v
    dense = dense.reshape(dims, order="F")
    # the same tensor, keyed in `seq`
    subs = np.unravel_index(k, dims, order="F")
    ks = np.zeros(n, dtype=np.int64)
    place = 1
    for m in seq:
        ks += subs[m] * place
        place *= dims[m]
    ck, cv = orc.contract_attract(ks, v, dims, seq, group, kind)
    want = np.einsum(spec, dense)
    out_dims = want.shape
    got = np.zeros(int(np.prod(out_dims)) if out_dims else 1)
    got[ck] = cv
    assert np.all(np.diff(ck) > 0)
    assert np.all(cv != 0)  # zero sums are dropped
    assert orc.allclose_fuzzy(got, np.ravel(want, order="F"), 1e-12)


def test_contract_attract_summation_order(orc):
    """The sums run in the reference's sort order: ascending diagonal value,
    and with two contracted indices the one whose last occurrence is later is
    the more significant.  The values make any other order give another sum
    (1e16 + 1 rounds to 1e16), and a sum of exactly zero is dropped."""
    # trace of a 3x3: diagonal 1e16, -1e16, 1 in that order -> 1; reversed -> 0
    k = np.array([0, 4, 8], dtype=np.int64)
    v = np.array([1e16, -1e16, 1.0])
    ck, cv = orc.contract_attract(k, v, (3, 3), (0, 1), (1, 1), (0,))
    assert ck.tolist() == [0] and cv.tolist() == [1.0]
    ck, cv = orc.contract_attract(k, v[::-1].copy(), (3, 3), (0, 1), (1, 1), (0,))
    assert len(ck) == 0  # 1 - 1e16 + 1e16 == 0: dropped
    # a(i,j,i,j), dims 2x2x2x2: j (last at 3) is more significant than i (last at 2)
    dims = (2, 2, 2, 2)
    key = lambda i, j: i + 2 * j + 4 * i + 8 * j
    k = np.array([key(0, 0), key(1, 0), key(0, 1)], dtype=np.int64)
    v = np.array([1e16, -1e16, 1.0])
    o = np.argsort(k)
    group, kind = _unary_groups("ijij")
    assert group == [1, 2, 1, 2]
    ck, cv = orc.contract_attract(k[o], v[o], dims, (0, 1, 2, 3), group, kind)
    assert cv.tolist() == [1.0]  # (i0,j0) + (i1,j0) + (i0,j1); i-major order would give 0


# ---------------------------------------------------------- sparse x dense
def test_sparse_dense_mult_known_answers(orc):
    """noLatticeMult with a dense B (SparseMIABase.h:1829-1910): C key
    r + l*R + t*L*R, one product a*b per A entry and B outer index, skipped
    when a or b is zero; checked against dense numpy on B modes that
    interleave outer and inter modes."""
    rng = np.random.default_rng(2)
    # B dims (2, 3, 4, 5): outer modes 0 and 2 (R = 8), inter modes 1 and 3 (T = 15)
    b_dims = (2, 3, 4, 5)
    b = rng.uniform(-1, 1, b_dims)
    b[rng.random(b_dims) < 0.3] = 0.0
    L = 7
    ak = np.sort(rng.choice(L * 15, 60, replace=False)).astype(np.int64)
    av = rng.uniform(-1, 1, len(ak))
    av[3] = 0.0  # a stored zero adds nothing
    ck, cv = orc.sparse_dense_mult(ak, av, L, b, b_dims, [0, 2], [1, 3])
    a = np.zeros((L, 15))
    a[ak % L, ak // L] = av
    bm = b.transpose(0, 2, 1, 3).reshape(8, 15, order="F")  # (r1 + 2 r2, t1 + 3 t2)
    want = np.einsum("lt,rt->rlt", a, bm)
    got = np.zeros(8 * L * 15)
    got[ck] = cv
    assert np.all(np.diff(ck) > 0)
    assert np.array_equal(got, np.ravel(want, order="F"))  # single products: exact
    assert set(ck.tolist()) == set(np.flatnonzero(np.ravel(want, order="F")).tolist())
    # no outer modes on B (pure element-wise) and no inter modes (outer product)
    ck, cv = orc.sparse_dense_mult(ak, av, L, bm[0].copy(), (15,), [], [0])
    assert np.array_equal(ck, np.sort(ck)) and len(ck) == np.count_nonzero(a * bm[0][None, :])
    ck, cv = orc.sparse_dense_mult(np.array([0, 2], np.int64), np.array([2.0, 3.0]), 3, np.array([1.0, 0.0, 5.0]), (3,),
                                   [0], [])
    assert ck.tolist() == [0, 2, 6, 8] and cv.tolist() == [2.0, 10.0, 3.0, 15.0]
