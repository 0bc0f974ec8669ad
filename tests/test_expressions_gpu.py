"""GPU checks of chained index expressions through the Python SparseMIA
mirror, after the reference's sparse_mia_mult_test.cpp:95-165: inner/
element-wise products, ``~`` between products with repeated element-wise
indices (``!!l``), ternary inner products, outer and element-wise products of
a tensor with itself, and complete inner products to a scalar.  Each result is
compared with dense numpy einsum (values within 1e-12, isEqualFuzzy)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rand_mia(orc, dims, nnz, seed):
    from libnt_amd.sparse import SparseMIA

    dom = int(np.prod(dims))
    k = orc.gen_distinct_keys(min(nnz, dom), dom, seed)  # unsorted, default order
    v = orc.gen_uniform_vals(len(k), seed + 1)
    return SparseMIA(dims, k, v)


def close(orc, got, want):
    return orc.allclose_fuzzy(np.ravel(got, order="F"), np.ravel(want, order="F"), TOL)


@pytest.fixture(scope="module")
def ops(orc):
    I, J, K, L, M, N = 3, 4, 5, 6, 3, 2
    return {
        "a": rand_mia(orc, (I, J, K, L), 250, 31),
        "b": rand_mia(orc, (J, L, M, N), 100, 33),
        "b2": rand_mia(orc, (J, L), 15, 35),
        "d": rand_mia(orc, (L,), 5, 37),
        "d2": rand_mia(orc, (J, L), 14, 39),
        "d2t": rand_mia(orc, (L, J), 13, 41),
        "sq": rand_mia(orc, (J, J), 9, 43),
        "e": rand_mia(orc, (I, L, J, K), 200, 45),
        "f": rand_mia(orc, (I, J, K, J), 150, 47),
    }


def test_inner_element_wise_products(ctx, orc, ops):
    """c2(i,j) = x(!i,!j,k,l) * y(k,l,!i,!j): inner product over k, l and
    element-wise over i, j (Inner/Element-Wise Products)."""
    x = ops["a"]
    y = rand_mia(orc, (5, 6, 3, 4), 120, 51)  # (k, l, i, j) sizes of x's (i, j, k, l) = (3,4,5,6)
    got = (x("!i!jkl") * y("kl!i!j")).assign("ij").to_dense()
    want = np.einsum("ijkl,klij->ij", x.to_dense(), y.to_dense())
    assert close(orc, got, want)


def test_repeated_element_wise_product(ctx, orc, ops):
    """c(i,j,k,l) = ~(a(i,!j,k,!!l) * b2(!j,!!l)) * d(!l)."""
    a, b2, d = ops["a"], ops["b2"], ops["d"]
    got = (~(a("i!jk!!l") * b2("!j!!l")) * d("!l")).assign("ijkl").to_dense()
    want = np.einsum("ijkl,jl,l->ijkl", a.to_dense(), b2.to_dense(), d.to_dense())
    assert close(orc, got, want)


@pytest.mark.parametrize("transposed", [False, True])
def test_ternary_inner_product(ctx, orc, ops, transposed):
    """c(i,k,m,n) = ~(a(i,!j,k,!l) * b(!j,!l,m,n)) * d2(j,l)  (or d2(l,j))."""
    a, b = ops["a"], ops["b"]
    first = ~(a("i!jk!l") * b("!j!lmn"))
    if transposed:
        got = (first * ops["d2t"]("lj")).assign("ikmn").to_dense()
        want = np.einsum("ijkl,jlmn,lj->ikmn", a.to_dense(), b.to_dense(), ops["d2t"].to_dense())
    else:
        got = (first * ops["d2"]("jl")).assign("ikmn").to_dense()
        want = np.einsum("ijkl,jlmn,jl->ikmn", a.to_dense(), b.to_dense(), ops["d2"].to_dense())
    assert close(orc, got, want)


def test_self_element_wise_product(ctx, orc, ops):
    """c3(i,j,l) = sq(i,!j) * sq(!j,l)  (self-multiplication)."""
    sq = ops["sq"]
    got = (sq("i!j") * sq("!jl")).assign("ijl").to_dense()
    want = np.einsum("ij,jl->ijl", sq.to_dense(), sq.to_dense())
    assert close(orc, got, want)


def test_complete_inner_products(ctx, orc, ops):
    a, e, f = ops["a"], ops["e"], ops["f"]
    s1 = a("ijkl") * e("iljk")
    w1 = np.einsum("ijkl,iljk->", a.to_dense(), e.to_dense())
    assert isinstance(s1, float) and orc.is_equal_fuzzy(s1, float(w1), TOL)
    s2 = f("ilkj") * f("ijkl")  # with itself, in two orders
    w2 = np.einsum("ilkj,ijkl->", f.to_dense(), f.to_dense())
    assert orc.is_equal_fuzzy(s2, float(w2), TOL)


def test_expression_rule_errors(ctx, orc, ops):
    a, b2 = ops["a"], ops["b2"]
    with pytest.raises(ValueError):  # one index, different element-wise counts
        a("i!jkl") * b2("jl")
    with pytest.raises(ValueError):
        a("i!jk!!l") * b2("!j!l")
    import libnt_amd

    with pytest.raises(libnt_amd.LntError):  # a contraction over modes of different sizes
        a("ijkj")
    # ~ on a plain atom changes nothing
    t = ~a("i!jkl")
    assert t.tokens == [("i", 0), ("j", 1), ("k", 0), ("l", 0)]
