"""GPU parity of the sparse comparison (lnt_compare: compare_with_sparse,
operator== and fuzzy_equals, SparseMIABase.h:1243-1296, 1616-1635) against
the CPU oracle's restatement of the reference's walk."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1.11e-16


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _both(ctx, orc, ak, av, bk, bv):
    """Exact and fuzzy (two precisions) results agree with the oracle, on the
    device and through host memory; returns the exact result."""
    out = None
    for prec in (None, 1e-6, 1e-12):
        mode = 0 if prec is None else 1
        want = orc.compare_sparse(ak, av, bk, bv, mode, prec or 1.0)
        got = ctx.compare(gpu(ak), gpu(av), gpu(bk), gpu(bv), prec)
        assert got == want, (prec, got, want)
        assert ctx.compare(np.ascontiguousarray(ak), np.ascontiguousarray(av), np.ascontiguousarray(bk),
                           np.ascontiguousarray(bv), prec) == want
        if prec is None:
            out = want
    return out


def _base(orc, n, seed, domain=1 << 40):
    k = np.sort(orc.gen_distinct_keys(n, domain, seed))
    v = orc.gen_uniform_vals(n, seed + 1) * 4.0  # |v| on both sides of 1
    v[v == 0] = 0.5
    return k, v


@pytest.mark.parametrize("n", [1, 7, 2047, 2048, 2049, 100_000])
def test_compare_edits(ctx, orc, n):
    k, v = _base(orc, n, n)
    rng = np.random.default_rng(n)
    assert _both(ctx, orc, k, v, k, v) is True
    # one value nudged: exact differs, fuzzy depends on size and precision
    for pos in {0, n // 2, n - 1, min(n - 1, 2047), min(n - 1, 2048)}:
        w = v.copy()
        w[pos] = w[pos] * (1 + 1e-9) + 1e-9
        assert _both(ctx, orc, k, v, k, w) is False
    # entries at or below the tolerance are absent, on either side
    for val in (0.0, -0.0, 1e-17, -TOL, TOL):
        extra = np.setdiff1d(np.arange(0, 4 * n + 8, dtype=np.int64), k)[: max(1, n // 10)]
        k2 = np.concatenate([k, extra])
        v2 = np.concatenate([v, np.full(len(extra), val)])
        o = np.argsort(k2)
        assert _both(ctx, orc, k, v, k2[o], v2[o]) is True
        assert _both(ctx, orc, k2[o], v2[o], k, v) is True
    # one significant extra, one missing key, one significant value made tiny
    extra = np.setdiff1d(np.arange(0, 4 * n + 8, dtype=np.int64), k)[:1]
    k2 = np.concatenate([k, extra])
    v2 = np.concatenate([v, [2e-16]])
    o = np.argsort(k2)
    assert _both(ctx, orc, k, v, k2[o], v2[o]) is False
    if n > 1:
        drop = int(rng.integers(0, n))
        assert _both(ctx, orc, k, v, np.delete(k, drop), np.delete(v, drop)) is False
        w = v.copy()
        w[drop] = 1e-20
        assert _both(ctx, orc, k, v, k, w) is False


def test_compare_random_against_oracle(ctx, orc):
    """Random mixtures of every edit on moderate sizes."""
    rng = np.random.default_rng(5)
    for trial in range(40):
        n = int(rng.integers(0, 5000))
        k, v = _base(orc, n, trial + 100, domain=1 << 16)
        bk, bv = k.copy(), v.copy()
        kind = trial % 5
        if kind == 1 and n:
            i = rng.integers(0, n, size=3)
            bv[i] = bv[i] * (1 + rng.choice([1e-13, 1e-8, 1e-3], size=3))
        elif kind == 2 and n:
            i = rng.integers(0, n, size=5)
            bv[i] = rng.choice([0.0, 1e-17, 5e-16], size=5)
        elif kind == 3 and n:
            keep = rng.random(n) > 0.01
            bk, bv = bk[keep], bv[keep]
        elif kind == 4:
            extra = np.setdiff1d(np.arange(1 << 16, dtype=np.int64), k)[: rng.integers(1, 50)]
            ev = rng.choice([0.0, 1e-18, 1.0], size=len(extra))
            bk = np.concatenate([bk, extra])
            bv = np.concatenate([bv, ev])
            o = np.argsort(bk)
            bk, bv = bk[o], bv[o]
        _both(ctx, orc, k, v, bk, bv)


def test_compare_special_values(ctx, orc):
    k = np.array([3, 8, 20], dtype=np.int64)
    v = np.array([1.0, np.nan, -2.5])
    assert _both(ctx, orc, k, v, k, v) is False  # NaN != NaN, as a == b
    v = np.array([1.0, -0.0, 5.0])
    w = np.array([1.0, 0.0, 5.0])
    assert _both(ctx, orc, k, v, k, w) is True
    empty_k, empty_v = np.zeros(0, dtype=np.int64), np.zeros(0)
    assert _both(ctx, orc, empty_k, empty_v, empty_k, empty_v) is True
    assert _both(ctx, orc, empty_k, empty_v, k[:1], np.array([1e-17])) is True
    assert _both(ctx, orc, k[:1], np.array([1.0]), empty_k, empty_v) is False


def test_front_end_equality(ctx, orc):
    """SparseMIA == and fuzzy_equals sort both operands into one order first."""
    from libnt_amd.sparse import SparseMIA

    dims = [6, 7, 8]
    keys = orc.gen_distinct_keys(150, 6 * 7 * 8, 3)
    vals = orc.gen_uniform_vals(150, 4)
    a = SparseMIA(dims, keys, vals, ctx=ctx)
    b = a.copy()
    b.sort([2, 0, 1])
    assert a == b
    c = a.copy()
    c.vals[5] += 1e-10
    assert not (a == c)
    assert a.fuzzy_equals(c, 1e-6)
    assert not (a == SparseMIA([6, 7, 9], keys, vals, ctx=ctx))


def test_compare_argument_errors(ctx):
    import libnt_amd

    k = gpu(np.array([1, 2], dtype=np.int64))
    v = gpu(np.array([1.0, 2.0]))
    with pytest.raises(libnt_amd.LntError):
        ctx.compare(k, v, k, v, 0.0)  # fuzzy precision must be positive
    with pytest.raises(libnt_amd.LntError):
        ctx.compare(k, v, k, v, -1.0)
