"""GPU parity of relinearisation, sort and change_sort_order (through the
C-ABI) against the CPU oracle: bit-exact keys and values."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def _bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


@pytest.mark.parametrize("domain", [1 << 20, 1 << 32, 12345, (1 << 61) + 3])
def test_testgen_matches_oracle(ctx, orc, domain):
    import torch

    n = min(100000, domain)
    k = torch.empty(n, dtype=torch.int64, device="cuda")
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.testgen_distinct_keys(k, domain, 77, offset=0)
    ctx.testgen_uniform(v, 78, offset=5)
    assert np.array_equal(host(k), orc.gen_distinct_keys(n, domain, 77))
    assert _bits_equal(host(v), orc.gen_uniform_vals(n, 78, offset=5))


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 1000, 4095, 4096, 4097, 10486, 65541, 1000003])
@pytest.mark.parametrize("bound", [1 << 20, (1 << 33) + 17, 1 << 62])
def test_sort_pairs_matches_oracle(ctx, orc, n, bound):
    print("sort_pairs n=%d bound=%d" % (n, bound), flush=True)
    keys = orc.gen_distinct_keys(n, bound, n ^ 0x55)
    vals = orc.gen_uniform_vals(n, n + 1)
    ek, ev = orc.radix_sort_pairs(keys, vals, bound)
    k, v = gpu(keys), gpu(vals)
    ctx.sort_pairs(k, v, bound)
    assert np.array_equal(host(k), ek)
    assert _bits_equal(host(v), ev)


@pytest.mark.parametrize("bound", [2, 3, 255, 256, 257, 65536, 1 << 24])
def test_sort_small_bounds(ctx, orc, bound):
    n = min(bound, 5000)
    keys = orc.gen_distinct_keys(n, bound, 3)
    vals = orc.gen_uniform_vals(n, 4)
    ek, ev = orc.radix_sort_pairs(keys, vals, bound)
    k, v = gpu(keys), gpu(vals)
    ctx.sort_pairs(k, v, bound)
    assert np.array_equal(host(k), ek) and _bits_equal(host(v), ev)


def test_sort_empty(ctx):
    import torch

    k = torch.empty(0, dtype=torch.int64, device="cuda")
    v = torch.empty(0, dtype=torch.float64, device="cuda")
    ctx.sort_pairs(k, v, 100)
    ctx.change_sort_order(k, v, [4, 5], [0, 1], [1, 0], True)


@pytest.mark.parametrize("case", range(12))
def test_change_linidx_matches_oracle(ctx, orc, case):
    rng = np.random.default_rng(case)
    order = int(rng.integers(1, 7))
    dims = [int(x) for x in rng.integers(1, 60, order)]
    total = int(np.prod(dims))
    n = min(total, 50000)
    keys = orc.gen_distinct_keys(n, total, case)
    cur = rng.permutation(order).astype(np.int32)
    new = rng.permutation(order).astype(np.int32)
    expect = orc.change_linidx(keys, dims, cur, new)
    k = gpu(keys)
    ctx.change_linidx(k, dims, cur, new)
    assert np.array_equal(host(k), expect)


def test_change_linidx_large_dims(ctx, orc):
    dims = [4096, 4096, 512, 64, 3]
    total = 4096 * 4096 * 512 * 64 * 3
    keys = orc.gen_distinct_keys(200000, total, 5)
    cur, new = [0, 1, 2, 3, 4], [3, 0, 4, 2, 1]
    k = gpu(keys)
    ctx.change_linidx(k, dims, cur, new)
    assert np.array_equal(host(k), orc.change_linidx(keys, dims, cur, new))


def _c1_operand(orc, seed):
    # 32^4 cells, 1% density: 10,486 distinct uniform coordinates, U[-1,1).
    keys = orc.gen_distinct_keys(10486, 1 << 20, seed)
    vals = orc.gen_uniform_vals(10486, seed + 1000)
    return keys, vals


@pytest.mark.parametrize("is_sorted", [False, True])
def test_change_sort_order_c1(ctx, orc, is_sorted):
    dims = [32, 32, 32, 32]
    ident = [0, 1, 2, 3]
    for seed in (1, 2):
        keys, vals = _c1_operand(orc, seed)
        if is_sorted:
            keys, vals = orc.radix_sort_pairs(keys, vals, 1 << 20)
        perm = orc.random_perm(4, seed + 500)
        ek, ev = orc.sort_seq(keys, vals, dims, ident, perm, is_sorted)
        k, v = gpu(keys), gpu(vals)
        ctx.change_sort_order(k, v, dims, ident, perm, is_sorted)
        assert np.array_equal(host(k), ek) and _bits_equal(host(v), ev)


@pytest.mark.parametrize("case", range(8))
def test_change_sort_order_various(ctx, orc, case):
    rng = np.random.default_rng(100 + case)
    order = int(rng.integers(2, 7))
    dims = [int(x) for x in rng.integers(2, 200, order)]
    total = int(np.prod(dims, dtype=object))
    while total >= (1 << 62):
        dims[int(np.argmax(dims))] //= 2
        total = int(np.prod(dims, dtype=object))
    n = int(min(total // 2, rng.integers(1, 300000)))
    keys = orc.gen_distinct_keys(n, total, case)
    vals = orc.gen_uniform_vals(n, case)
    cur = rng.permutation(order).astype(np.int32)
    new = rng.permutation(order).astype(np.int32)
    ek, ev = orc.sort_seq(keys, vals, dims, cur, new, False)
    k, v = gpu(keys), gpu(vals)
    ctx.change_sort_order(k, v, dims, cur, new, False)
    assert np.array_equal(host(k), ek) and _bits_equal(host(v), ev)


SHARED_PREFIX_CASES = [
    # dims, cur, new: the fastest modes of `new` keep their order in `cur`
    ([256, 256, 256, 256], [0, 1, 2, 3], [0, 2, 1, 3]),      # 2^16: two passes fewer
    ([8, 8, 8, 8], [0, 1, 2, 3], [1, 3, 0, 2]),              # modes 1, 3 in order: 6 bits
    ([12, 5, 7, 9], [0, 1, 2, 3], [0, 2, 1, 3]),              # 12 = 4 * 3: two low bits
    ([8, 16, 3, 50], [0, 1, 3, 2], [0, 1, 2, 3]),             # 8 * 16 = 2^7
    ([6, 4, 10], [2, 0, 1], [2, 1, 0]),                       # shared mode of size 10: one bit
    ([1 << 20, 1 << 20, 16], [0, 1, 2], [0, 2, 1]),           # 20 ordered bits of 44
    ([3, 5, 7], [0, 1, 2], [0, 2, 1]),                        # odd product: nothing skipped
    ([1 << 30, 2, 2], [0, 1, 2], [0, 2, 1]),                  # two bits left to sort
    ([4, 4, 4, 4, 4], [1, 0, 2, 3, 4], [1, 0, 4, 3, 2]),      # two shared modes, 2^4
]


@pytest.mark.parametrize("case", range(len(SHARED_PREFIX_CASES)))
def test_change_sort_order_sorted_shared_prefix(ctx, orc, case):
    """A sorted tensor whose new sequence starts with modes in their old
    relative order: the radix passes skip the low bits whose order the input
    carries, and the result equals the oracle's full sort."""
    dims, cur, new = SHARED_PREFIX_CASES[case]
    total = int(np.prod(dims, dtype=object))
    for n in (1, 5000, 200000):
        n = min(n, total)
        keys = orc.gen_distinct_keys(n, total, case * 7 + n)
        vals = orc.gen_uniform_vals(n, case + n)
        order = np.argsort(keys, kind="stable")
        keys, vals = keys[order], vals[order]  # sorted in the current sequence
        ek, ev = orc.sort_seq(keys, vals, dims, cur, new, False)
        k, v = gpu(keys), gpu(vals)
        ctx.change_sort_order(k, v, dims, cur, new, True)
        assert np.array_equal(host(k), ek), (case, n)
        assert _bits_equal(host(v), ev), (case, n)


def test_change_sort_order_sorted_shared_prefix_large(ctx):
    """20M sorted keys over 256^4 into [0, 2, 1, 3] (two radix passes)
    against a numpy closed form, then back into the identity sequence."""
    import torch

    n, perm = 20_000_000, [0, 2, 1, 3]
    k = torch.empty(n, dtype=torch.int64, device="cuda")
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.testgen_distinct_keys(k, 1 << 32, 17)
    ctx.testgen_uniform(v, 18)
    ctx.sort_pairs(k, v, 1 << 32)
    k0, v0 = host(k), host(v)
    nk = _relin_bytes_np(k0, perm)
    order = np.argsort(nk, kind="stable")
    ctx.change_sort_order(k, v, [256] * 4, [0, 1, 2, 3], perm, True)
    assert np.array_equal(host(k), nk[order])
    assert _bits_equal(host(v), v0[order])
    ctx.change_sort_order(k, v, [256] * 4, perm, [0, 1, 2, 3], True)
    assert np.array_equal(host(k), k0)
    assert _bits_equal(host(v), v0)


def test_sort_repeat_is_bitwise_identical(ctx, orc):
    """Look-back or ranking races would show up as run-to-run differences."""
    import torch

    n = 3_000_000
    k0 = torch.empty(n, dtype=torch.int64, device="cuda")
    v0 = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.testgen_distinct_keys(k0, 1 << 40, 9)
    ctx.testgen_uniform(v0, 10)
    outs = []
    for _ in range(3):
        k, v = k0.clone(), v0.clone()
        ctx.sort_pairs(k, v, 1 << 40)
        outs.append((k, v))
    torch.cuda.synchronize()
    for k, v in outs[1:]:
        assert torch.equal(k, outs[0][0]) and torch.equal(v.view(torch.int64), outs[0][1].view(torch.int64))
    assert bool((outs[0][0][1:] > outs[0][0][:-1]).all())


def test_key_range_error_leaves_input_untouched(ctx, orc):
    import libnt_amd

    keys = orc.gen_distinct_keys(5000, 1 << 20, 1)
    keys[1234] = 1 << 20  # one key outside [0, 32^4)
    vals = orc.gen_uniform_vals(5000, 2)
    k, v = gpu(keys), gpu(vals)
    with pytest.raises(libnt_amd.LntError) as e:
        ctx.change_sort_order(k, v, [32, 32, 32, 32], [0, 1, 2, 3], [1, 0, 3, 2], False)
    assert e.value.status == 5
    assert np.array_equal(host(k), keys) and _bits_equal(host(v), vals)
    keys[1234] = -7
    k = gpu(keys)
    with pytest.raises(libnt_amd.LntError):
        ctx.sort_pairs(k, v, 1 << 20)
    assert np.array_equal(host(k), keys)


def test_invalid_arguments(ctx):
    import libnt_amd
    import torch

    k = torch.zeros(4, dtype=torch.int64, device="cuda")
    v = torch.zeros(4, dtype=torch.float64, device="cuda")
    for dims, cur, new in (([2, 2], [0, 0], [0, 1]), ([2, 0], [0, 1], [1, 0]), ([2, 2], [0, 2], [1, 0])):
        with pytest.raises(libnt_amd.LntError) as e:
            ctx.change_sort_order(k, v, dims, cur, new, False)
        assert e.value.status == 1
    with pytest.raises(libnt_amd.LntError) as e:
        ctx.change_sort_order(k, v, [1 << 32, 1 << 32], [0, 1], [1, 0], False)
    assert e.value.status == 6


def test_host_memory_mode(ctx, orc):
    keys, vals = _c1_operand(orc, 7)
    dims, ident, perm = [32, 32, 32, 32], [0, 1, 2, 3], [3, 1, 0, 2]
    ek, ev = orc.sort_seq(keys, vals, dims, ident, perm, False)
    k, v = keys.copy(), vals.copy()
    ctx.change_sort_order(k, v, dims, ident, perm, False)
    assert np.array_equal(k, ek) and _bits_equal(v, ev)
    k2 = keys.copy()
    ctx.change_linidx(k2, dims, ident, perm)
    assert np.array_equal(k2, orc.change_linidx(keys, dims, ident, perm))


def test_c2_scale_change_sort_order_against_torch_sort(ctx):
    """50M keys over 256^4: the result equals torch.sort (an independent
    library sort) of the relinearised keys, with values following."""
    import torch

    n = 50_000_000
    dims, ident, perm = [256] * 4, [0, 1, 2, 3], [2, 0, 3, 1]
    k = torch.empty(n, dtype=torch.int64, device="cuda")
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.testgen_distinct_keys(k, 1 << 32, 11)
    ctx.testgen_uniform(v, 12)
    ctx.sort_pairs(k, v, 1 << 32)  # sorted in the default order
    rk = k.clone()
    ctx.change_linidx(rk, dims, ident, perm)
    ref_k, idx = torch.sort(rk)
    ref_v = v[idx]
    ctx.change_sort_order(k, v, dims, ident, perm, True)
    torch.cuda.synchronize()
    assert torch.equal(k, ref_k)
    assert torch.equal(v.view(torch.int64), ref_v.view(torch.int64))


def _relin_bytes_np(keys, perm):
    """Closed form for dims 256^4: byte t of the new key is byte perm[t] of the old."""
    out = np.zeros_like(keys)
    for t, m in enumerate(perm):
        out |= ((keys >> (8 * m)) & 255) << (8 * t)
    return out


@pytest.mark.parametrize("n", [3_000_000, 20_000_000])
def test_change_linidx_large_n_against_closed_form(ctx, n):
    """Relinearisation over millions of keys (every block of the grid and every
    grid-stride round) against a numpy closed form, repeated to catch
    intermittent errors."""
    import torch

    perm = [2, 0, 3, 1]
    k = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.testgen_distinct_keys(k, 1 << 32, 11)
    expect = _relin_bytes_np(host(k), perm)
    for _ in range(3):
        rk = k.clone()
        ctx.change_linidx(rk, [256] * 4, [0, 1, 2, 3], perm)
        assert np.array_equal(host(rk), expect)


def test_change_sort_order_large_n_against_closed_form(ctx):
    import torch

    n, perm = 20_000_000, [2, 0, 3, 1]
    k = torch.empty(n, dtype=torch.int64, device="cuda")
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.testgen_distinct_keys(k, 1 << 32, 13)
    ctx.testgen_uniform(v, 14)
    nk = _relin_bytes_np(host(k), perm)
    order = np.argsort(nk, kind="stable")
    ek, ev = nk[order], host(v)[order]
    ctx.change_sort_order(k, v, [256] * 4, [0, 1, 2, 3], perm, False)
    assert np.array_equal(host(k), ek)
    assert _bits_equal(host(v), ev)


@pytest.mark.parametrize("block,items", [(256, 8), (256, 12), (256, 16), (512, 8), (512, 12), (512, 16),
                                         (1024, 8), (1024, 12)])
def test_sort_configurations_match_oracle(ctx, orc, block, items):
    """Every onesweep configuration (threads per block x items per thread,
    digit widths 8 and 6, narrow keys on and off) gives the oracle's result,
    for sizes around one and several tiles."""
    ctx.set_param("sort_block", block)
    ctx.set_param("sort_items", items)
    try:
        tile = block * items
        for n in (1, tile - 1, tile, tile + 1, 3 * tile + 17, 250000):
            for bound in (1 << 20, (1 << 40) + 3):
                for digits in (8, 6):
                    for narrow in (1, 0):
                        ctx.set_param("digit_bits", digits)
                        ctx.set_param("narrow_keys", narrow)
                        keys = orc.gen_distinct_keys(min(n, bound), bound, n + digits)
                        vals = orc.gen_uniform_vals(len(keys), n)
                        ek, ev = orc.radix_sort_pairs(keys, vals, bound)
                        k, v = gpu(keys), gpu(vals)
                        ctx.sort_pairs(k, v, bound)
                        assert np.array_equal(host(k), ek), (n, bound, digits, narrow)
                        assert _bits_equal(host(v), ev)
        dims, ident, perm = [64, 32, 128, 16], [0, 1, 2, 3], [2, 0, 3, 1]
        keys = orc.gen_distinct_keys(300000, 64 * 32 * 128 * 16, 5)
        vals = orc.gen_uniform_vals(300000, 6)
        ek, ev = orc.sort_seq(keys, vals, dims, ident, perm, False)
        k, v = gpu(keys), gpu(vals)
        ctx.change_sort_order(k, v, dims, ident, perm, False)
        assert np.array_equal(host(k), ek) and _bits_equal(host(v), ev)
    finally:
        ctx.set_param("sort_block", 1024)
        ctx.set_param("sort_items", 12)
        ctx.set_param("digit_bits", 8)
        ctx.set_param("narrow_keys", 1)


@pytest.mark.parametrize("hist_blocks", [1, 63, 64, 65, 1000, 4096])
def test_histogram_block_counts(ctx, orc, hist_blocks):
    """The digit histogram's workgroup count (1..4096; the partial rows are
    reduced in groups of 64, so counts off a multiple of 64 leave a ragged
    last group) never changes the result: a full sort and a sorted-input
    change_sort_order, each against the oracle."""
    print("histogram blocks %d" % hist_blocks, flush=True)
    ctx.set_param("hist_blocks", hist_blocks)
    try:
        n, bound = 700001, (1 << 40) + 7
        keys = orc.gen_distinct_keys(n, bound, 5)
        vals = orc.gen_uniform_vals(n, 6)
        ek, ev = orc.radix_sort_pairs(keys, vals, bound)
        k, v = gpu(keys), gpu(vals)
        ctx.sort_pairs(k, v, bound)
        assert np.array_equal(host(k), ek) and _bits_equal(host(v), ev)
        dims, ident, perm = [64, 64, 64, 64], [0, 1, 2, 3], [2, 0, 3, 1]
        keys = np.sort(orc.gen_distinct_keys(300000, 1 << 24, 7))
        vals = orc.gen_uniform_vals(300000, 8)
        ek, ev = orc.sort_seq(keys, vals, dims, ident, perm, True)
        k, v = gpu(keys), gpu(vals)
        ctx.change_sort_order(k, v, dims, ident, perm, True)
        assert np.array_equal(host(k), ek) and _bits_equal(host(v), ev)
    finally:
        ctx.set_param("hist_blocks", 4096)


@pytest.mark.parametrize("seed", range(6))
def test_change_sort_order_random_sweep(ctx, orc, seed):
    """Random orders 1-6, dims 1..300 (powers of two or not), random current
    and new mode sequences, sorted or unsorted input, up to 200k entries."""
    rng = np.random.default_rng([seed, 99])
    for trial in range(5):
        order = int(rng.integers(1, 7))
        dims = [int(x) for x in rng.choice([1, 2, 3, 7, 16, 31, 64, 100, 128, 256, 300], order)]
        dom = int(np.prod(dims))
        n = int(min(dom, rng.integers(1, 200000)))
        cur = [int(x) for x in rng.permutation(order)]
        new = [int(x) for x in rng.permutation(order)]
        is_sorted = bool(rng.integers(0, 2))
        print("sweep seed %d trial %d dims %s cur %s new %s n %d sorted %d" % (seed, trial, dims, cur, new, n, is_sorted),
              flush=True)
        keys = orc.gen_distinct_keys(n, dom, 1000 * seed + trial)
        if is_sorted:
            keys = np.sort(keys)
        vals = orc.gen_uniform_vals(n, 7 + trial)
        ek, ev = orc.sort_seq(keys, vals, dims, cur, new, is_sorted)
        k, v = gpu(keys), gpu(vals)
        ctx.change_sort_order(k, v, dims, cur, new, is_sorted)
        assert np.array_equal(host(k), ek) and _bits_equal(host(v), ev)
