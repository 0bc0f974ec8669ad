"""Operands of 2^32 entries and more (the reference's size_t lengths,
LibMIARadix.h:444, FunctionUtil.h:373-419).  The inputs are closed forms, so
the expected result is known entry by entry without a CPU pass over 4G keys;
every comparison runs on the GPU in slices."""
import pytest

pytestmark = pytest.mark.gpu

N = (1 << 31) + 3  # A and B together hold 2^32 + 6 entries
SLICE = 1 << 28


def _need_gib(gib):
    import torch

    free, _ = torch.cuda.mem_get_info(0)
    if free < gib * (1 << 30):
        pytest.skip(f"needs {gib} GiB of free device memory")


def _check_slices(ck, cv, n, key_of, val_of):
    import torch

    assert ck.shape[0] >= n
    for s in range(0, n, SLICE):
        e = min(n, s + SLICE)
        k = key_of(torch.arange(s, e, device=ck.device, dtype=torch.int64))
        assert torch.equal(ck[s:e], k), (s, e)
        assert torch.equal(cv[s:e].view(torch.int64), val_of(k).view(torch.int64)), (s, e)


def test_merge_of_2_pow_32_entries(ctx):
    """A = 0..N-1 with value k, B = 1..N with value k/2.  The merge path of
    these inputs puts an equal pair across every even diagonal, so the first
    chunk boundary (diagonal 2^31) must move to keep key 2^30 whole: a split
    pair shows up as a duplicate key and a count of N + 2."""
    import torch

    import libnt_amd

    _need_gib(150)

    dev = torch.device("cuda:0")
    ak = torch.arange(0, N, device=dev, dtype=torch.int64)
    av = ak.to(torch.float64)
    bk = torch.arange(1, N + 1, device=dev, dtype=torch.int64)
    bv = bk.to(torch.float64) * 0.5
    torch.cuda.synchronize()
    print(f"merge ADD of {N} + {N} entries", flush=True)
    ck = torch.empty(2 * N, device=dev, dtype=torch.int64)
    cv = torch.empty(2 * N, device=dev, dtype=torch.float64)
    n = ctx.merge_into(ak, av, bk, bv, libnt_amd.OP_ADD, ck, cv)
    assert n == N + 1

    def add(k):
        v = torch.where(k < N, k.to(torch.float64), 0.0)
        return v + torch.where(k >= 1, k.to(torch.float64) * 0.5, 0.0)

    _check_slices(ck, cv, n, lambda i: i, add)
    print(f"merge HADAMARD of {N} + {N} entries", flush=True)
    n = ctx.merge_into(ak, av, bk, bv, libnt_amd.OP_HADAMARD, ck[:N], cv[:N])
    assert n == N - 1
    _check_slices(ck, cv, n, lambda i: i + 1, lambda k: k.to(torch.float64) * (k.to(torch.float64) * 0.5))
    del ck, cv, ak, av, bk, bv
    torch.cuda.empty_cache()
    ctx.trim()


M = (1 << 32) + 6  # entries of the sort test
MUL, ADD = 2654435761, 12345  # MUL is prime, so i -> (MUL i + ADD) mod M is a bijection


def _perm(i):
    """(MUL * i + ADD) mod M in int64 without overflow (i < 2^33)."""
    x = (i * (MUL >> 16)) % M
    x = (x * 65536) % M
    return (x + i * (MUL & 0xFFFF) + ADD) % M


def _check_pairs(keys, vals, key_to_k):
    """Every key appears exactly twice, keys ascend, each entry's value is
    the input position of an entry with that key, and the two entries of a
    key keep their input order (stability)."""
    import torch

    n = keys.shape[0]
    prev = None
    for s in range(0, n, SLICE):
        e = min(n, s + SLICE)
        k = keys[s:e]
        v = vals[s:e]
        assert torch.equal(k[0::2], k[1::2]), s
        assert bool((k[2::2] > k[1:-1:2]).all()), s
        if prev is not None:
            assert int(k[0]) > prev, s
        prev = int(k[-1])
        assert bool((v[0::2] < v[1::2]).all()), s
        orig = key_to_k(k)
        assert bool((orig >= 0).all()) and bool((orig < M // 2).all()), s
        assert torch.equal(_perm(v.to(torch.int64)) >> 1, orig), s


def test_sort_of_2_pow_32_entries(ctx):
    """2^32 + 6 entries: key perm(i) >> 1 (every key in [0, M/2) twice, in
    scattered positions), value i.  lnt_sort_pairs must give keys j >> 1
    with each key's two values ascending; then change_sort_order over dims
    (65536, 32769) to the sequence [1, 0] (relinearised in the chunk sorts)
    must give the same pairs, stably, in the new order; lnt_change_linidx
    back to [0, 1] must then give the original keys."""
    import torch

    _need_gib(180)
    dev = torch.device("cuda:0")
    keys = torch.empty(M, device=dev, dtype=torch.int64)
    vals = torch.empty(M, device=dev, dtype=torch.float64)
    for s in range(0, M, SLICE):
        i = torch.arange(s, min(M, s + SLICE), device=dev, dtype=torch.int64)
        keys[s:s + i.shape[0]] = _perm(i) >> 1
        vals[s:s + i.shape[0]] = i.to(torch.float64)
    torch.cuda.synchronize()
    print(f"sort_pairs of {M} entries", flush=True)
    ctx.sort_pairs(keys, vals, M // 2)
    for s in range(0, M, SLICE):
        e = min(M, s + SLICE)
        assert torch.equal(keys[s:e], torch.arange(s, e, device=dev, dtype=torch.int64) >> 1), s
    _check_pairs(keys, vals, lambda k: k)

    dims, d0 = [65536, 32769], 65536
    print(f"change_sort_order of {M} entries", flush=True)
    ctx.change_sort_order(keys, vals, dims, [0, 1], [1, 0], True)
    _check_pairs(keys, vals, lambda k: k // dims[1] + d0 * (k % dims[1]))
    print(f"change_linidx of {M} entries", flush=True)
    expect = [None] * ((M + SLICE - 1) // SLICE)
    for t, s in enumerate(range(0, M, SLICE)):
        k = keys[s:s + SLICE]
        expect[t] = k // dims[1] + d0 * (k % dims[1])
    ctx.change_linidx(keys, dims, [1, 0], [0, 1])
    for t, s in enumerate(range(0, M, SLICE)):
        assert torch.equal(keys[s:s + SLICE], expect[t]), s
    del keys, vals
    torch.cuda.empty_cache()
    ctx.trim()
