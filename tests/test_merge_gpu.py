"""GPU parity of the sparse merges (+, -, Hadamard) through the C-ABI against
the CPU oracle: bit-exact keys and values."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = (0, 1, 2)


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def _bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def _pair(orc, na, nb, overlap, domain, seed):
    keys = orc.gen_distinct_keys(na + nb - overlap, domain, seed)
    ak = np.sort(keys[:na])
    bk = np.sort(np.concatenate([keys[:overlap], keys[na:]]))
    av = orc.gen_uniform_vals(na, seed + 1)
    bv = orc.gen_uniform_vals(len(bk), seed + 2)
    return ak, av, bk, bv


def _check(ctx, orc, ak, av, bk, bv):
    for op in OPS:
        ek, ev = orc.merge(ak, av, bk, bv, op)
        ck, cv = ctx.merge(gpu(ak), gpu(av), gpu(bk), gpu(bv), op)
        assert np.array_equal(host(ck), ek), op
        assert _bits_equal(host(cv), ev), op


@pytest.mark.parametrize(
    "na,nb,ov",
    [(0, 0, 0), (0, 7, 0), (7, 0, 0), (1, 1, 1), (1, 1, 0), (1024, 1024, 1024), (2048, 2048, 2048),
     (3000, 5000, 1500), (5000, 3000, 3000), (4097, 1, 1), (1, 4097, 1), (100000, 100000, 50000),
     (250000, 10, 5), (333333, 444444, 111111)],
)
def test_merge_matches_oracle(ctx, orc, na, nb, ov):
    print("merge na=%d nb=%d overlap=%d" % (na, nb, ov), flush=True)
    ak, av, bk, bv = _pair(orc, na, nb, ov, 1 << 40, na * 7 + nb)
    _check(ctx, orc, ak, av, bk, bv)


def test_merge_dense_interleaving(ctx, orc):
    # Every key of a small domain present in A or B or both: equal keys fall on
    # every possible tile and thread boundary.
    rng = np.random.default_rng(3)
    for trial in range(4):
        dom = 20000 + trial * 777
        where = rng.integers(0, 3, dom)  # 0: A only, 1: B only, 2: both
        ak = np.flatnonzero(where != 1).astype(np.int64)
        bk = np.flatnonzero(where != 0).astype(np.int64)
        av = rng.uniform(-1, 1, len(ak))
        bv = rng.uniform(-1, 1, len(bk))
        _check(ctx, orc, ak, av, bk, bv)


@pytest.mark.parametrize("narrow", [1, 0])
@pytest.mark.parametrize("block,items", [(256, 8), (256, 12), (256, 16), (512, 8), (512, 12), (512, 16),
                                         (1024, 8), (1024, 12), (1024, 16)])
def test_merge_tile_shapes(ctx, orc, block, items, narrow):
    """Every merge tile shape, with tile-relative 32-bit LDS keys allowed or
    every tile forced onto the 64-bit path, gives the oracle's result: dense
    interleaving (equal keys on every tile and thread boundary), long runs of
    one operand, a tile-sized operand against a single key, and keys on both
    sides of 2^32."""
    print("merge tile %dx%d narrow %d" % (block, items, narrow), flush=True)
    ctx.set_param("merge_block", block)
    ctx.set_param("merge_items", items)
    ctx.set_param("narrow_keys", narrow)
    try:
        for top in ((1 << 32) - 1, 1 << 32):
            rng = np.random.default_rng(top % 97)
            where = rng.integers(0, 3, 3 * block * items)
            ak = top - np.flatnonzero(where != 1)[::-1].astype(np.int64) * 5
            bk = top - np.flatnonzero(where != 0)[::-1].astype(np.int64) * 5
            _check(ctx, orc, ak, rng.uniform(-1, 1, len(ak)), bk, rng.uniform(-1, 1, len(bk)))
            _check(ctx, orc, ak[:-1], rng.uniform(-1, 1, len(ak) - 1), bk, rng.uniform(-1, 1, len(bk)))
        rng = np.random.default_rng(block + items)
        where = rng.integers(0, 3, 50000)
        ak = np.flatnonzero(where != 1).astype(np.int64)
        bk = np.flatnonzero(where != 0).astype(np.int64)
        _check(ctx, orc, ak, rng.uniform(-1, 1, len(ak)), bk, rng.uniform(-1, 1, len(bk)))
        ak = np.arange(0, 3 * block * items, 3, dtype=np.int64)
        bk = np.unique(np.concatenate([np.arange(1, 60000, 7), [ak[-1], 10 ** 6]])).astype(np.int64)
        _check(ctx, orc, ak, rng.uniform(-1, 1, len(ak)), bk, rng.uniform(-1, 1, len(bk)))
        ak, av, bk, bv = _pair(orc, block * items, 1, 1, 1 << 40, 11)
        _check(ctx, orc, ak, av, bk, bv)
        _check(ctx, orc, bk, bv, ak, av)
    finally:
        ctx.set_param("merge_block", 512)
        ctx.set_param("merge_items", 0)
        ctx.set_param("narrow_keys", 1)


def test_merge_mixed_tile_widths(ctx, orc):
    """One merge whose tiles are partly narrow (keys within 2^32 - 1 of the
    tile's smallest) and partly wide (keys spread over 2^62), with the
    switches between them on tile and thread boundaries, equal keys on both
    sides, and the B key after a tile far above it."""
    rng = np.random.default_rng(77)
    parts = [np.arange(0, 30000, dtype=np.int64),                            # dense, narrow
             rng.integers(1 << 40, 1 << 62, 20000, dtype=np.int64),          # sparse, wide
             (1 << 62) + np.arange(0, 40000, 3, dtype=np.int64),             # dense again
             np.array([(1 << 62) + (1 << 33), (1 << 63) - 1], dtype=np.int64)]  # far keys
    dom = np.unique(np.concatenate(parts))
    where = rng.integers(0, 3, len(dom))
    ak, bk = dom[where != 1], dom[where != 0]
    _check(ctx, orc, ak, rng.uniform(-1, 1, len(ak)), bk, rng.uniform(-1, 1, len(bk)))
    _check(ctx, orc, bk, rng.uniform(-1, 1, len(bk)), ak[: len(ak) // 2], rng.uniform(-1, 1, len(ak) // 2))
    # a single far B key after a narrow A run: the peeked key is no match
    ak = np.arange(0, 9000, dtype=np.int64)
    bk = np.array([5, 8999, (1 << 32) + 5, (1 << 32) + 8999], dtype=np.int64)
    _check(ctx, orc, ak, rng.uniform(-1, 1, len(ak)), bk, rng.uniform(-1, 1, len(bk)))


@pytest.mark.parametrize("seed", range(4))
def test_merge_random_sweep(ctx, orc, seed):
    """Random sizes (0..300k), overlaps and key domains from 2^10 to 2^62, so
    that tiles are narrow, wide or mixed, every operation against the oracle."""
    rng = np.random.default_rng([seed, 7])
    for trial in range(6):
        dom = 1 << int(rng.integers(10, 63))
        na, nb = (int(rng.integers(0, min(300000, dom // 2) + 1)) for _ in range(2))
        ov = int(rng.integers(0, min(na, nb) + 1))
        print("merge sweep seed %d trial %d na %d nb %d ov %d dom 2^%d" % (seed, trial, na, nb, ov, dom.bit_length() - 1),
              flush=True)
        ak, av, bk, bv = _pair(orc, na, nb, ov, dom, 100 * seed + trial)
        _check(ctx, orc, ak, av, bk, bv)


def test_merge_signed_zeros(ctx, orc):
    ak = np.array([1, 4], dtype=np.int64)
    av = np.array([-0.0, 0.0])
    bk = np.array([2, 3, 4], dtype=np.int64)
    bv = np.array([-0.0, 0.0, -0.0])
    _check(ctx, orc, ak, av, bk, bv)


def test_merge_c2_scaled(ctx, orc):
    """c2 layout at 1/10 scale: 7.5M distinct keys over 2^32, A = first 5M,
    B = first 2.5M + last 2.5M."""
    import torch

    n_shared, n_only = 2_500_000, 2_500_000
    total = n_shared + 2 * n_only
    ak = torch.empty(n_shared + n_only, dtype=torch.int64, device="cuda")
    bk = torch.empty(n_shared + n_only, dtype=torch.int64, device="cuda")
    ctx.testgen_distinct_keys(ak, 1 << 32, 1, offset=0)
    ctx.testgen_distinct_keys(bk[:n_shared], 1 << 32, 1, offset=0)
    ctx.testgen_distinct_keys(bk[n_shared:], 1 << 32, 1, offset=n_shared + n_only)
    av = torch.empty(len(ak), dtype=torch.float64, device="cuda")
    bv = torch.empty(len(bk), dtype=torch.float64, device="cuda")
    ctx.testgen_uniform(av, 2)
    ctx.testgen_uniform(bv, 3)
    ctx.sort_pairs(ak, av, 1 << 32)
    ctx.sort_pairs(bk, bv, 1 << 32)
    hak, hav, hbk, hbv = host(ak), host(av), host(bk), host(bv)
    for op in OPS:
        ck, cv = ctx.merge(ak, av, bk, bv, op)
        ek, ev = orc.merge(hak, hav, hbk, hbv, op)
        assert len(ek) == (n_shared if op == 2 else total)
        assert np.array_equal(host(ck), ek) and _bits_equal(host(cv), ev)


def test_merge_library_allocates(ctx, orc):
    import libnt_amd

    L = libnt_amd.lib()
    ak, av, bk, bv = _pair(orc, 3000, 4000, 1000, 1 << 30, 5)
    ga, gav, gb, gbv = gpu(ak), gpu(av), gpu(bk), gpu(bv)
    ck, cv = ctypes.c_void_p(), ctypes.c_void_p()
    nc = ctypes.c_int64()
    libnt_amd.check(L.lnt_merge(ctx.handle, ga.data_ptr(), gav.data_ptr(), len(ak), gb.data_ptr(),
                                gbv.data_ptr(), len(bk), 0, ctypes.byref(ck), ctypes.byref(cv),
                                ctypes.byref(nc), libnt_amd.MEM_DEVICE))
    out_k = np.empty(nc.value, dtype=np.int64)
    out_v = np.empty(nc.value, dtype=np.float64)
    import torch

    torch.cuda.synchronize()
    tk = torch.empty(nc.value, dtype=torch.int64, device="cuda")
    tv = torch.empty(nc.value, dtype=torch.float64, device="cuda")
    # copy the library-owned buffers into tensors through a device-to-device memcpy
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(ctypes.c_void_p(tk.data_ptr()), ck, nc.value * 8, 3) == 0
    assert hip.hipMemcpy(ctypes.c_void_p(tv.data_ptr()), cv, nc.value * 8, 3) == 0
    out_k, out_v = host(tk), host(tv)
    ek, ev = orc.merge(ak, av, bk, bv, 0)
    assert np.array_equal(out_k, ek) and _bits_equal(out_v, ev)
    libnt_amd.check(L.lnt_free(ctx.handle, ck, libnt_amd.MEM_DEVICE))
    libnt_amd.check(L.lnt_free(ctx.handle, cv, libnt_amd.MEM_DEVICE))


def test_merge_host_memory(ctx, orc):
    ak, av, bk, bv = _pair(orc, 10000, 8000, 3000, 1 << 33, 8)
    for op in OPS:
        ck, cv = ctx.merge(ak, av, bk, bv, op)
        ek, ev = orc.merge(ak, av, bk, bv, op)
        assert np.array_equal(ck, ek) and _bits_equal(cv, ev)


def test_merge_capacity_and_sortedness_errors(ctx, orc):
    import libnt_amd
    import torch

    ak, av, bk, bv = _pair(orc, 100, 100, 10, 1 << 20, 9)
    small = torch.empty(10, dtype=torch.int64, device="cuda")
    smallv = torch.empty(10, dtype=torch.float64, device="cuda")
    with pytest.raises(libnt_amd.LntError) as e:
        ctx.merge_into(gpu(ak), gpu(av), gpu(bk), gpu(bv), 0, small, smallv)
    assert e.value.status == 1
    bad = ak.copy()
    bad[[10, 11]] = bad[[11, 10]]
    ctx.set_validation(2)
    try:
        with pytest.raises(libnt_amd.LntError) as e:
            ctx.merge(gpu(bad), gpu(av), gpu(bk), gpu(bv), 0)
        assert e.value.status == 4
    finally:
        ctx.set_validation(1)


# ------------------------------------------------ sortMerge (unsorted operands)
def _unsorted_pair_with_zeros(orc, na, nb, overlap, domain, seed, zero_frac):
    keys = orc.gen_distinct_keys(na + nb - overlap, domain, seed)
    ak = keys[:na].copy()
    bk = np.concatenate([keys[:overlap], keys[na:]])
    rng = np.random.default_rng(seed)
    bk = rng.permutation(bk)
    av = orc.gen_uniform_vals(na, seed + 1)
    bv = orc.gen_uniform_vals(len(bk), seed + 2)
    av[rng.random(na) < zero_frac] = 0.0
    bv[rng.random(len(bk)) < zero_frac] = -0.0
    return ak, av, bk, bv


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("na,nb,ov,zf", [(1, 1, 1, 1.0), (1000, 700, 300, 0.3), (50000, 80000, 20000, 0.05),
                                         (3, 0, 0, 0.5), (0, 5, 0, 0.5), (200000, 200000, 200000, 0.5)])
def test_sort_merge_matches_oracle(ctx, orc, op, na, nb, ov, zf):
    """Unsorted operands with planted zeros (+0 and -0): keys whose entries are
    all zero vanish, the smallest key always stays (SparseMIA.h:628-665)."""
    ak, av, bk, bv = _unsorted_pair_with_zeros(orc, na, nb, ov, 1 << 30, na + 7 * nb, zf)
    ek, ev = orc.sort_merge(ak, av, bk, bv, op, 1 << 30)
    ck, cv = ctx.sort_merge(gpu(ak), gpu(av), gpu(bk), gpu(bv), op, 1 << 30)
    assert np.array_equal(host(ck), ek)
    assert _bits_equal(host(cv), ev)
    k2, v2 = ctx.sort_merge(ak, av, bk, bv, op, 1 << 30)  # host memory
    assert np.array_equal(k2, ek) and _bits_equal(v2, ev)


def test_sort_merge_zero_rules(ctx, orc):
    """Hand-made cases of collect_duplicates' rule: A={1:0, 3:1, 5:2, 7:0},
    B={1:2, 4:0, 5:-2, 7:0}.  Key 1 is the smallest: kept (0 + 2); 4 (only a
    zero) and 7 (zeros only) vanish; 5 sums to an explicit 0 and stays."""
    ak = np.array([3, 1, 7, 5], dtype=np.int64)
    av = np.array([1.0, 0.0, 0.0, 2.0])
    bk = np.array([5, 1, 7, 4], dtype=np.int64)
    bv = np.array([-2.0, 2.0, -0.0, 0.0])
    ck, cv = ctx.sort_merge(gpu(ak), gpu(av), gpu(bk), gpu(bv), 0, 8)
    assert host(ck).tolist() == [1, 3, 5] and host(cv).tolist() == [2.0, 1.0, 0.0]
    # the smallest key stays even when all its entries are zero
    ck, cv = ctx.sort_merge(gpu(np.array([0, 6])), gpu(np.array([0.0, 3.0])), gpu(np.array([0])),
                            gpu(np.array([0.0])), 0, 8)
    assert host(ck).tolist() == [0, 6] and host(cv).tolist() == [0.0, 3.0]
    # subtraction negates B (-1*b): 3 - 3 = 0 kept, B-only 4 -> -5
    ck, cv = ctx.sort_merge(gpu(np.array([2, 3])), gpu(np.array([1.0, 3.0])), gpu(np.array([4, 3])),
                            gpu(np.array([5.0, 3.0])), 1, 8)
    assert host(ck).tolist() == [2, 3, 4] and host(cv).tolist() == [1.0, 0.0, -5.0]


def test_front_end_sort_merge_sequence(orc):
    """C = A(i,j,k) + B(k,i,j) with neither sorted: the smaller operand is
    relinearised into the other's sequence, C is sorted in that sequence
    (SparseMIA.h:836-862), and C equals dense arithmetic."""
    from libnt_amd.sparse import SparseMIA

    dims_a, dims_b = [7, 6, 5], [5, 7, 6]
    for na, nb in ((150, 60), (40, 170)):
        ak = orc.gen_distinct_keys(na, 210, na)
        bk = orc.gen_distinct_keys(nb, 210, nb)
        A = SparseMIA(dims_a, ak, orc.gen_uniform_vals(na, 1), seq=[2, 0, 1])
        B = SparseMIA(dims_b, bk, orc.gen_uniform_vals(nb, 2), seq=[1, 2, 0])
        Ad, Bd = A.to_dense(), B.to_dense()
        C = A("ijk") + B("kij")
        # B's sequence [1,2,0] over (k,i,j) is (i,j,k): A's modes [0,1,2]
        want_seq = [2, 0, 1] if na > nb else [0, 1, 2]
        assert C.seq == want_seq and C.is_sorted
        assert np.array_equal(C.to_dense(), Ad + Bd.transpose(1, 2, 0))


def test_front_end_destructive_merge(orc):
    """A("ijk") += B("kij") and -= write the merge back into A
    (SparseMIA::merge, SparseMIA.h:1030-1050), sorted or not."""
    from libnt_amd.sparse import SparseMIA

    for a_sorted, b_sorted in ((True, True), (False, False), (True, False)):
        ka = orc.gen_distinct_keys(80, 5 * 6 * 7, 41)
        kb = orc.gen_distinct_keys(90, 7 * 5 * 6, 42)
        va, vb = orc.gen_uniform_vals(80, 43), orc.gen_uniform_vals(90, 44)
        A = SparseMIA([5, 6, 7], ka, va)
        B = SparseMIA([7, 5, 6], kb, vb)
        if a_sorted:
            A.sort()
        if b_sorted:
            B.sort()
        da = np.zeros(5 * 6 * 7)
        da[ka] = va
        db = np.zeros(7 * 5 * 6)
        db[kb] = vb
        x = A("ijk")
        x += B("kij")
        A.sort([0, 1, 2])
        got = np.zeros(5 * 6 * 7)
        got[A.keys.cpu().numpy()] = A.vals.cpu().numpy()
        want = np.zeros(5 * 6 * 7)
        for i in range(5):
            for j in range(6):
                for k in range(7):
                    want[i + 5 * (j + 6 * k)] = da[i + 5 * (j + 6 * k)] + db[k + 7 * (i + 5 * j)]
        assert np.array_equal(got, want), (a_sorted, b_sorted)
        y = A("ijk")
        y -= A.copy()("ijk")
        assert not np.any(A.vals.cpu().numpy())
