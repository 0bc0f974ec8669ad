"""GPU parity of sparse x dense products with no contracted index
(lnt_sparse_dense_mult: SparseMIABase::noLatticeMult with a DenseMIA,
SparseMIABase.h:1829-1910) against the CPU oracle: keys and values bit-exact
(each value is one product), through the C-ABI on device and host memory and
through the Python front end."""
import numpy as np
import pytest

import libnt_amd

pytestmark = pytest.mark.gpu


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


CASES = [
    # (L, b_dims, r_modes, i_modes, nnz A, zero fraction of B)
    (7, (2, 3, 4, 5), [0, 2], [1, 3], 60, 0.3),
    (1, (1000,), [], [0], 300, 0.5),            # pure element-wise
    (50, (13,), [0], [], 40, 0.2),              # pure outer product
    (256, (64, 32, 16), [2, 0], [1], 5000, 0.1),
    (1 << 12, (300, 257), [1], [0], 200_000, 0.0),
    (3, (1 << 20,), [0], [], 3, 0.0),           # a long dense row per A entry: 3M outputs
]


@pytest.mark.parametrize("L,b_dims,r_modes,i_modes,nnz,zf", CASES)
def test_sparse_dense_matches_oracle(ctx, orc, L, b_dims, r_modes, i_modes, nnz, zf):
    print("sparse x dense", L, b_dims, r_modes, i_modes, nnz, flush=True)
    rng = np.random.default_rng(nnz)
    T = int(np.prod([b_dims[m] for m in i_modes])) if i_modes else 1
    ak = np.sort(orc.gen_distinct_keys(min(nnz, L * T), L * T, 3))
    av = orc.gen_uniform_vals(len(ak), 4)
    av[:: max(1, len(av) // 7)] = 0.0  # stored zeros add nothing
    b = orc.gen_uniform_vals(int(np.prod(b_dims)), 5)
    b[rng.random(len(b)) < zf] = 0.0
    ek, ev = orc.sparse_dense_mult(ak, av, L, b.reshape(b_dims, order="F"), b_dims, r_modes, i_modes)
    ck, cv = ctx.sparse_dense_mult(gpu(ak), gpu(av), L, gpu(b), b_dims, r_modes, i_modes)
    assert np.array_equal(ck.cpu().numpy(), ek)
    assert np.array_equal(bits(cv.cpu().numpy()), bits(ev))
    hk, hv = ctx.sparse_dense_mult(ak, av, L, b, b_dims, r_modes, i_modes)
    assert np.array_equal(hk, ek) and np.array_equal(bits(hv), bits(ev))


def test_sparse_dense_errors(ctx):
    ak, av = gpu(np.array([0, 5], np.int64)), gpu(np.array([1.0, 2.0]))
    b = gpu(np.ones(6))
    with pytest.raises(libnt_amd.LntError) as e:  # key 5 -> t = 2 >= T = 2 with L = 2
        ctx.sparse_dense_mult(ak, av, 2, b, (3, 2), [0], [1])
    assert e.value.status == 5
    with pytest.raises(libnt_amd.LntError):  # mode 1 listed twice
        ctx.sparse_dense_mult(ak, av, 3, b, (3, 2), [1], [1])
    ck, cv = ctx.sparse_dense_mult(gpu(np.empty(0, np.int64)), gpu(np.empty(0)), 3, b, (3, 2), [0], [1])
    assert len(ck) == 0


def test_sparse_dense_front_end(ctx, orc):
    """A(i,j,!k) * B(!k,l) and A(i,!j,!k) * B(!k,!j) through the SparseMIA
    mirror, against numpy."""
    from libnt_amd.sparse import DenseMIA, SparseMIA

    dims = (4, 5, 6)
    k = orc.gen_distinct_keys(60, 120, 7)
    a = SparseMIA(dims, k, orc.gen_uniform_vals(60, 8))
    rng = np.random.default_rng(9)
    bd = rng.uniform(-1, 1, (6, 3))
    bd[rng.random(bd.shape) < 0.25] = 0.0
    be = rng.uniform(-1, 1, (6, 5))
    da = a.to_dense()
    c = (a("ij!k") * DenseMIA.from_array(bd)("!kl")).assign("ijkl")
    want = np.einsum("ijk,kl->ijkl", da, bd)
    assert np.array_equal(c.to_dense(), want)
    assert c.nnz == np.count_nonzero(want)
    c2 = (a("i!j!k") * DenseMIA.from_array(be)("!k!j")).assign("ikj")
    assert np.array_equal(c2.to_dense(), np.einsum("ijk,kj->ikj", da, be))
    with pytest.raises(ValueError):  # contracted index
        a("ijk") * DenseMIA.from_array(bd)("kl")


@pytest.mark.parametrize("seed", range(3))
def test_sparse_dense_random_sweep(ctx, orc, seed):
    """Random dense operands of order 1-3 (each mode an outer or an inter
    mode of the product, in random order), random outer sizes of A and
    densities of both operands, against the oracle bit for bit."""
    rng = np.random.default_rng([seed, 11])
    for trial in range(6):
        nbm = int(rng.integers(1, 4))
        b_dims = tuple(int(x) for x in rng.integers(1, 12, nbm))
        modes = [int(m) for m in rng.permutation(nbm)]
        cut = int(rng.integers(0, nbm + 1))
        r_modes, i_modes = modes[:cut], modes[cut:]
        L = int(rng.integers(1, 40))
        T = int(np.prod([b_dims[m] for m in i_modes])) if i_modes else 1
        nnz = int(rng.integers(1, L * T + 1))
        print("dense sweep seed %d L %d b_dims %s r %s i %s nnz %d" % (seed, L, b_dims, r_modes, i_modes, nnz), flush=True)
        ak = np.sort(orc.gen_distinct_keys(nnz, L * T, 10 * seed + trial))
        av = orc.gen_uniform_vals(len(ak), 4 + trial)
        av[rng.random(len(av)) < 0.1] = 0.0
        b = orc.gen_uniform_vals(int(np.prod(b_dims)), 5 + trial)
        b[rng.random(len(b)) < 0.3] = 0.0
        ek, ev = orc.sparse_dense_mult(ak, av, L, b.reshape(b_dims, order="F"), b_dims, r_modes, i_modes)
        ck, cv = ctx.sparse_dense_mult(gpu(ak), gpu(av), L, gpu(b), b_dims, r_modes, i_modes)
        assert np.array_equal(ck.cpu().numpy(), ek)
        assert np.array_equal(bits(cv.cpu().numpy()), bits(ev))
