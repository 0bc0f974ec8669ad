"""GPU parity of unary contraction/attraction (lnt_contract_attract:
SparseMIABase::contract_attract, SparseMIABase.h:1913-2053) against the CPU
oracle's restatement of the reference's walk: keys and values bit-exact (the
sums run in the reference's order), through the C-ABI on device and host
memory and through the Python front end."""
import numpy as np
import pytest

import libnt_amd
from unary_cases import UNARY_CASES, unary_groups

pytestmark = pytest.mark.gpu


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def _operand(orc, dims, n, seed, seq):
    """n distinct entries of a tensor with dims, keyed in mode order seq, in
    generator (unsorted) order."""
    dom = int(np.prod(dims))
    k = orc.gen_distinct_keys(min(n, dom), dom, seed)
    v = orc.gen_uniform_vals(len(k), seed + 1) * 5
    subs = np.unravel_index(k, dims, order="F")
    ks = np.zeros(len(k), dtype=np.int64)
    place = 1
    for m in seq:
        ks += subs[m] * place
        place *= dims[m]
    return ks, v


def _check(ctx, orc, k, v, dims, seq, group, kind):
    ek, ev = orc.contract_attract(k, v, dims, seq, group, kind)
    ck, cv = ctx.contract_attract(gpu(k), gpu(v), dims, seq, group, kind)
    assert np.array_equal(ck.cpu().numpy(), ek)
    assert np.array_equal(bits(cv.cpu().numpy()), bits(ev))
    hk, hv = ctx.contract_attract(np.ascontiguousarray(k), np.ascontiguousarray(v), dims, seq, group, kind)
    assert np.array_equal(hk, ek) and np.array_equal(bits(hv), bits(ev))
    return ek, ev


@pytest.mark.parametrize("labels,dims,spec", UNARY_CASES)
def test_unary_matches_oracle(ctx, orc, labels, dims, spec):
    print("unary", labels, dims, flush=True)
    group, kind = unary_groups(labels)
    seq = list(np.random.default_rng(len(labels) + 7).permutation(len(dims)))
    k, v = _operand(orc, dims, int(np.prod(dims)) // 3, 5, seq)
    _check(ctx, orc, k, v, dims, seq, group, kind)


@pytest.mark.parametrize("labels,dims", [
    ("ijiklm", (40, 30, 40, 20, 10, 8)),        # 7.7M cells: long sums over i
    ("IjIkLL", (64, 50, 64, 30, 16, 16)),       # two attractions
    ("ijij", (1000, 300, 1000, 300)),           # 2 groups, 9e10 cells, sparse diagonals
    ("ii", (1 << 20, 1 << 20)),                 # trace of a 2^20 x 2^20 matrix
])
def test_unary_large(ctx, orc, labels, dims):
    """Bigger operands (1M entries), a tenth of them planted on the diagonals
    so that the sums are not empty."""
    print("unary large", labels, dims, flush=True)
    group, kind = unary_groups(labels)
    seq = list(range(len(dims)))[::-1]
    n = 1_000_000
    k, v = _operand(orc, dims, n, 11, seq)
    # plant entries on every diagonal: copy the subscript of each group's first mode
    rng = np.random.default_rng(3)
    sel = rng.choice(len(k), len(k) // 10, replace=False)
    subs = list(np.unravel_index(np.zeros(len(k), np.int64), dims, order="F"))
    place = [int(np.prod([dims[m] for m in seq[:t]])) for t in range(len(seq))]
    for t, m in enumerate(seq):
        subs[m] = (k // place[t]) % dims[m]
    for g in set(group) - {0}:
        modes = [m for m in range(len(dims)) if group[m] == g]
        for m in modes[1:]:
            subs[m][sel] = subs[modes[0]][sel]
    ks = np.zeros(len(k), dtype=np.int64)
    for t, m in enumerate(seq):
        ks += subs[m] * place[t]
    ks, first = np.unique(ks, return_index=True)
    v = v[first]
    ek, ev = _check(ctx, orc, ks, v, dims, seq, group, kind)
    assert len(ek) > 0


def test_unary_summation_order_and_zero_drop(ctx, orc):
    """Values whose sum depends on the order: the reference's order gives 1,
    others 0; a sum of zero is dropped."""
    k = np.array([0, 4, 8], dtype=np.int64)
    for v, want in ((np.array([1e16, -1e16, 1.0]), [1.0]), (np.array([1.0, -1e16, 1e16]), [])):
        ek, ev = _check(ctx, orc, k, v, (3, 3), (0, 1), (1, 1), (0,))
        assert ev.tolist() == want
    key = lambda i, j: i + 2 * j + 4 * i + 8 * j
    k = np.array([key(0, 0), key(1, 0), key(0, 1)], dtype=np.int64)
    ek, ev = _check(ctx, orc, k, np.array([1e16, -1e16, 1.0]), (2, 2, 2, 2), (0, 1, 2, 3), [1, 2, 1, 2], [0, 0])
    assert ev.tolist() == [1.0]


def test_unary_edges(ctx, orc):
    e = np.empty(0, np.int64)
    ck, cv = ctx.contract_attract(gpu(e), gpu(np.empty(0)), (4, 4), (0, 1), (1, 1), (0,))
    assert len(ck) == 0
    # no groups: the tensor in its default order, sorted, zeros dropped
    k = np.array([5, 1, 7, 2], dtype=np.int64)  # in seq (1, 0) of dims (2, 4)
    v = np.array([1.0, 0.0, -2.0, 3.0])
    _check(ctx, orc, k, v, (2, 4), (1, 0), (0, 0), [])
    # every mode contracted: one entry at key 0
    _check(ctx, orc, np.array([0, 3, 6, 8], np.int64), np.array([1.0, 5.0, 7.0, 2.0]), (3, 3), (0, 1), (1, 1), (0,))
    with pytest.raises(libnt_amd.LntError) as err:  # group modes of different sizes
        ctx.contract_attract(gpu(k), gpu(v), (2, 4), (0, 1), (1, 1), (0,))
    assert err.value.status == 7
    with pytest.raises(libnt_amd.LntError):  # group number out of range
        ctx.contract_attract(gpu(k), gpu(v), (2, 4), (0, 1), (1, 2), (0,))


def test_unary_front_end(ctx, orc):
    """The Python front end: repeated labels contract or attract at once, and
    the result takes part in products like any tensor."""
    from libnt_amd.sparse import SparseMIA

    dims = (5, 6, 5, 6, 5, 3)
    k, v = _operand(orc, dims, 1500, 21, list(range(6)))
    a = SparseMIA(dims, k, v)
    dense = a.to_dense()
    for labels, spec, out in (("ijiklm", "ijiklm->jklm", "jklm"), ("ijijim", "ijijim->m", "m"),
                              ("i!jk!jlm", "ijkjlm->iklmj", "iklmj")):
        got = a(labels).assign(out).to_dense()
        assert orc.allclose_fuzzy(np.ravel(got, order="F"), np.ravel(np.einsum(spec, dense), order="F"), 1e-12)
    s = SparseMIA((7, 7), *_operand(orc, (7, 7), 20, 23, [0, 1]))("ii")
    assert isinstance(s, float)
    bb = SparseMIA((6, 5, 3, 4), *_operand(orc, (6, 5, 3, 4), 200, 25, [0, 1, 2, 3]))
    got = (a("ijiklm") * bb("klmn")).assign("jn").to_dense()
    want = np.einsum("ijiklm,klmn->jn", dense, bb.to_dense())
    assert orc.allclose_fuzzy(np.ravel(got, order="F"), np.ravel(want, order="F"), 1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_unary_random_sweep(ctx, orc, seed):
    """Random index strings of order 2-6 over a four-letter alphabet (every
    repeated letter contracted or, at random, attracted), random sizes and
    mode sequences, up to 5000 entries, against the oracle bit for bit."""
    rng = np.random.default_rng([seed, 5])
    done = 0
    while done < 6:
        order = int(rng.integers(2, 7))
        labels = "".join(rng.choice(list("abcd"), order))
        if all(labels.count(c) == 1 for c in labels):
            continue  # no repeated index: not a unary contraction
        upper = {c for c in set(labels) if labels.count(c) > 1 and rng.random() < 0.5}
        labels = "".join(c.upper() if c in upper else c for c in labels)
        size = {c: int(rng.integers(2, 9)) for c in set(labels.lower())}
        dims = tuple(size[c.lower()] for c in labels)
        group, kind = unary_groups(labels)
        seq = [int(x) for x in rng.permutation(order)]
        n = int(min(np.prod(dims), rng.integers(1, 5000)))
        print("unary sweep seed %d labels %s dims %s seq %s n %d" % (seed, labels, dims, seq, n), flush=True)
        k, v = _operand(orc, dims, n, 100 * seed + done, seq)
        _check(ctx, orc, k, v, dims, seq, group, kind)
        done += 1
