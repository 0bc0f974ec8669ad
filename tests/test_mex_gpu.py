"""GPU parity of the mex-compatible entry points (include/lnt_mex.h) against
the CPU oracle, with the mex conventions: 1-based uint8 mode sequences,
in-place mutation of the caller's arrays (SparseSortMex, SparseChangeLinIdxMex,
SparsePermuteMex) and new result arrays (SparseMergeMex, SparseLatticeMultMex),
the mex's early returns and argument checks (src/MATLAB/mex/*.cpp,
check_sparse.cpp)."""
import numpy as np
import pytest

import libnt_amd
from libnt_amd import mex

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def _tensor_sorted(orc, dims, n, seed):
    """n distinct keys in the default (1..order) linearisation, sorted, with
    values: a tensor as the toolbox holds it after a sort."""
    dom = int(np.prod(dims))
    n = min(n, dom)
    k = np.sort(orc.gen_distinct_keys(n, dom, seed))
    v = orc.gen_uniform_vals(n, seed + 1)
    return k, v


# ----------------------------------------------------------------- sorting --
@pytest.mark.parametrize("n", [1, 2, 5, 2999, 3000, 100_003, 1 << 20])
def test_sparse_sort_matches_introsort(ctx, orc, n):
    print("mex sort n", n, flush=True)
    k = orc.gen_distinct_keys(n, 1 << 41, n)  # unsorted, unique
    v = orc.gen_uniform_vals(n, n + 7)
    ek, ev = orc.introsort_pairs(k, v)
    kk, vv = k.copy(), v.copy()
    mex.sparse_sort(ctx, vv, kk)
    assert np.array_equal(kk, ek)
    assert np.array_equal(bits(vv), bits(ev))


@pytest.mark.parametrize("n,distinct", [(60, 7), (3_000, 300), (8_192, 5000), (8_193, 40), (50_000, 5000),
                                        (65_536, 100)])
def test_sparse_sort_payload_bits_and_duplicates(ctx, orc, n, distinct):
    """Values only move: NaN payloads, signed zeros and any 8-byte class passed
    as double come back bit for bit.  With equal indices the data lands in the
    order the reference's Introsort + DualSwapper leaves it (the oracle's
    swap-for-swap restatement), in LDS (n <= 8192) and in global memory."""
    rng = np.random.default_rng(n)
    k = rng.integers(0, distinct, n).astype(np.int64)  # many duplicates
    raw = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)  # arbitrary bit patterns
    v = raw.view(np.float64).copy()
    kk, vv = k.copy(), v.copy()
    print("mex sort with duplicates", n, distinct, flush=True)
    mex.sparse_sort(ctx, vv, kk)
    ek, ev = orc.introsort_pairs(k, v)
    assert np.array_equal(kk, ek)
    assert np.array_equal(bits(vv), bits(ev))
    # the reference's order of equal keys is not the stable one
    if n >= 3000:
        assert not np.array_equal(bits(vv), raw[np.argsort(k, kind="stable")])


def test_sparse_sort_duplicates_beyond_replay(ctx):
    """Above 65,536 entries the Introsort is not replayed: indices sorted,
    equal indices in their input order."""
    rng = np.random.default_rng(8)
    n = 70_000
    k = rng.integers(0, 700, n).astype(np.int64)
    raw = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    kk, vv = k.copy(), raw.view(np.float64).copy()
    mex.sparse_sort(ctx, vv, kk)
    o = np.argsort(k, kind="stable")
    assert np.array_equal(kk, k[o]) and np.array_equal(bits(vv), raw[o])


def test_sparse_sort_edges(ctx):
    k = np.empty(0, np.int64)
    v = np.empty(0, np.float64)
    mex.sparse_sort(ctx, v, k)  # empty: nothing to do
    k = np.array([5, 0, (1 << 62) + 3, 7], np.int64)
    v = np.array([1.0, 2.0, 3.0, 4.0])
    mex.sparse_sort(ctx, v, k)
    assert k.tolist() == [0, 5, 7, (1 << 62) + 3] and v.tolist() == [2.0, 1.0, 4.0, 3.0]
    k = np.array([3, -1, 2], np.int64)
    v = np.zeros(3)
    with pytest.raises(libnt_amd.LntError) as e:
        mex.sparse_sort(ctx, v, k)
    assert e.value.status == 5  # LNT_ERR_KEY_RANGE
    assert k.tolist() == [3, -1, 2]  # nothing written on error


# ------------------------------------------------------ relinearisation --
RELIN_CASES = [
    ((7, 5, 3), [1, 2, 3], [3, 1, 2]),
    ((64, 32, 16, 8), [1, 2, 3, 4], [1, 3, 2, 4]),
    ((64, 32, 16, 8), [2, 1, 4, 3], [4, 3, 2, 1]),
    ((3, 5, 7, 11, 2), [1, 2, 3, 4, 5], [5, 4, 3, 2, 1]),
    ((1024, 1024), [1, 2], [2, 1]),
]


@pytest.mark.parametrize("dims,cur,new", RELIN_CASES)
def test_sparse_change_linidx(ctx, orc, dims, cur, new):
    print("mex change_linidx", dims, cur, new, flush=True)
    k = orc.gen_distinct_keys(min(200_000, int(np.prod(dims))), int(np.prod(dims)), 11)
    want = orc.change_linidx(k, dims, np.array(cur) - 1, np.array(new) - 1)
    got = k.copy()
    mex.sparse_change_linidx(ctx, got, np.array(cur, np.uint8), np.array(new, np.uint8), dims)
    assert np.array_equal(got, want)
    # and back
    mex.sparse_change_linidx(ctx, got, np.array(new, np.uint8), np.array(cur, np.uint8), dims)
    assert np.array_equal(got, k)


def test_sparse_change_linidx_early_returns(ctx):
    """Equal sequences or no indices return before any check of the
    sequences (sparse_change_linIdx.cpp:30-31, 49-50)."""
    k = np.array([10**15, 3], np.int64)  # out of range for dims: untouched
    mex.sparse_change_linidx(ctx, k, [1, 1], [1, 1], [2, 2])
    assert k.tolist() == [10**15, 3]
    mex.sparse_change_linidx(ctx, np.empty(0, np.int64), [0, 9], [3, 3], [2, 2])
    with pytest.raises(libnt_amd.LntError):  # 0 is no 1-based mode number
        mex.sparse_change_linidx(ctx, np.array([1], np.int64), [0, 1], [1, 2], [2, 2])
    with pytest.raises(libnt_amd.LntError):  # not a permutation
        mex.sparse_change_linidx(ctx, np.array([1], np.int64), [1, 1], [1, 2], [2, 2])


PERMUTE_CASES = [
    ((64, 32, 16, 8), [1, 2, 3, 4], [1, 3, 2, 4]),  # shared fast prefix: skipped low bits
    ((64, 32, 16, 8), [1, 2, 3, 4], [4, 3, 2, 1]),
    ((64, 32, 16, 8), [3, 1, 2, 4], [1, 2, 4, 3]),
    ((3, 5, 7, 11, 2), [1, 2, 3, 4, 5], [1, 2, 5, 3, 4]),  # odd dims: no bits skipped
    ((1000, 999, 3), [2, 1, 3], [3, 2, 1]),
    ((1 << 12, 1 << 12, 1 << 6), [1, 2, 3], [2, 1, 3]),
]


@pytest.mark.parametrize("dims,cur,new", PERMUTE_CASES)
@pytest.mark.parametrize("n", [1, 777, 300_000])
def test_sparse_permute(ctx, orc, dims, cur, new, n):
    """SparseNT.permute's pair of calls: SparseChangeLinIdxMex then
    SparsePermuteMex on a tensor sorted in cur, against the oracle's sort of
    the relinearised tensor."""
    print("mex permute", dims, cur, new, n, flush=True)
    cur0, new0 = np.array(cur) - 1, np.array(new) - 1
    # a tensor sorted in the cur linearisation
    k, v = _tensor_sorted(orc, dims, n, 21)
    ek, ev = orc.sort_seq(k, v, dims, cur0, new0, True)
    kk, vv = k.copy(), v.copy()
    mex.sparse_change_linidx(ctx, kk, np.array(cur, np.uint8), np.array(new, np.uint8), dims)
    mex.sparse_permute(ctx, vv, kk, np.array(cur, np.uint8), np.array(new, np.uint8), dims)
    assert np.array_equal(kk, ek)
    assert np.array_equal(bits(vv), bits(ev))


def test_sparse_permute_early_returns_and_range(ctx, orc):
    k = np.array([5, 1, 3], np.int64)  # unsorted, left alone when the orders agree
    v = np.array([1.0, 2.0, 3.0])
    mex.sparse_permute(ctx, v, k, [2, 1], [2, 1], [4, 4])
    assert k.tolist() == [5, 1, 3]
    with pytest.raises(libnt_amd.LntError) as e:  # index beyond prod(dims)
        mex.sparse_permute(ctx, v, np.array([1, 2, 16], np.int64), [1, 2], [2, 1], [4, 4])
    assert e.value.status == 5


# -------------------------------------------------------------------- merge --
def _operands(orc, na, nb, seed, domain=1 << 34):
    ak = np.sort(orc.gen_distinct_keys(na, domain, seed)) if na else np.empty(0, np.int64)
    bk = np.sort(orc.gen_distinct_keys(nb, domain, seed + 100)) if nb else np.empty(0, np.int64)
    av = orc.gen_uniform_vals(na, seed + 1) if na else np.empty(0)
    bv = orc.gen_uniform_vals(nb, seed + 2) if nb else np.empty(0)
    return ak, av, bk, bv


@pytest.mark.parametrize("na,nb", [(1, 1), (1000, 3), (70_001, 65_536), (1 << 20, 1 << 19)])
@pytest.mark.parametrize("op", [0, 1, 0.75, 1.5])
def test_sparse_merge(ctx, orc, na, nb, op):
    print("mex merge", na, nb, op, flush=True)
    ak, av, bk, bv = _operands(orc, na, nb, na + nb)
    # some shared keys
    m = min(na, nb) // 3
    bk = np.unique(np.concatenate([bk[: nb - m], ak[:m]]))
    bv = orc.gen_uniform_vals(len(bk), 9)
    ek, ev = orc.mex_merge(ak, av, bk, bv, int(np.floor(op)))
    cv, ck = mex.sparse_merge(ctx, av, ak, bv, bk, op)
    assert np.array_equal(ck, ek)
    assert np.array_equal(bits(cv), bits(ev))


def test_sparse_merge_empty_operands(ctx, orc):
    """The mex's own shortcuts (sparse_merge.cpp:26-57): signed zeros follow
    -1*b for a subtraction from nothing, a copy otherwise."""
    k = np.array([0, 4, 9, 12], np.int64)
    v = np.array([0.0, -0.0, 2.5, -3.0])
    e_k, e_v = np.empty(0, np.int64), np.empty(0)
    for op in (0, 1):
        for (ak, av, bk, bv) in ((k, v, e_k, e_v), (e_k, e_v, k, v), (e_k, e_v, e_k, e_v)):
            wk, wv = orc.mex_merge(ak, av, bk, bv, op)
            cv, ck = mex.sparse_merge(ctx, av, ak, bv, bk, op)
            assert np.array_equal(ck, wk) and np.array_equal(bits(cv), bits(wv)), (op, len(ak), len(bk))
    cv, _ = mex.sparse_merge(ctx, e_v, e_k, v, k, 1)
    assert np.array_equal(bits(cv), bits(np.array([-0.0, 0.0, -2.5, 3.0])))


def test_sparse_merge_op_checks(ctx):
    k = np.array([1], np.int64)
    v = np.array([1.0])
    for op in (-0.5, 2.0, float("nan")):
        with pytest.raises(libnt_amd.LntError):
            mex.sparse_merge(ctx, v, k, v, k, op)


# ---------------------------------------------------------- lattice product --
LATTICE_CASES = [
    # (ah, aw, d, bw, nA, nB)
    (1024, 1024, 1, 1024, 10486, 10486),
    (50, 40, 3, 30, 1500, 900),
    (1 << 20, 64, 1, 4096, 200000, 3000),
    (7, 5, 2, 9, 30, 40),
]


@pytest.mark.parametrize("case", LATTICE_CASES)
@pytest.mark.parametrize("algorithm", [0, 1, 2, 3, 4, 5, 5.9, -0.5])
def test_sparselattice_mult(ctx, orc, case, algorithm):
    ah, aw, d, bw, n_a, n_b = case
    print("mex lattice", case, "algorithm", algorithm, flush=True)
    ak = np.sort(orc.gen_distinct_keys(min(n_a, ah * aw * d), ah * aw * d, 3))
    av = orc.gen_uniform_vals(len(ak), 4)
    bk = np.sort(orc.gen_distinct_keys(min(n_b, aw * bw * d), aw * bw * d, 5))
    bv = orc.gen_uniform_vals(len(bk), 6)
    ek, ev = orc.lattice_mult(ak, av, (ah, aw, d), bk, bv, (aw, bw, d))
    cv, ck = mex.sparselattice_mult(ctx, av, ak, (ah, aw, d), bv, bk, (aw, bw, d), algorithm)
    assert np.array_equal(ck, ek)
    assert orc.allclose_fuzzy(cv, ev, 1e-12)
    # Every algorithm's GPU counterpart sums each row in ascending B row
    # order: the bits of csc_times's sparse accumulator (and of mtt_times's,
    # which performs the same additions; algorithm 4 returns mtt_times's
    # result through the mex's fall-through into case 5).
    method = {0: "spa", 1: "spa", 2: "spa", 3: "spa", 4: "mtt", 5: "mtt"}[int(algorithm) if algorithm > 0 else 0]
    sk, sv = orc.lattice_mult_by(method, ak, av, (ah, aw, d), bk, bv, (aw, bw, d))
    assert np.array_equal(ck, sk) and np.array_equal(cv.view(np.int64), sv.view(np.int64))


def test_sparselattice_mult_checks_and_empty(ctx, orc):
    k = np.array([0, 3], np.int64)
    v = np.array([1.0, 2.0])
    for alg in (6, -1, float("nan")):
        with pytest.raises(libnt_amd.LntError):
            mex.sparselattice_mult(ctx, v, k, (2, 2, 1), v, k, (2, 2, 1), alg)
    with pytest.raises(libnt_amd.LntError) as e:  # A width != B height
        mex.sparselattice_mult(ctx, v, k, (2, 2, 1), v, k, (3, 2, 1), 0)
    assert e.value.status == 7
    # disjoint inner indices: empty product, empty outputs
    cv, ck = mex.sparselattice_mult(ctx, np.array([1.0]), np.array([0], np.int64), (2, 2, 1),
                                    np.array([1.0]), np.array([1], np.int64), (2, 2, 1), 0)
    assert len(cv) == 0 and len(ck) == 0
