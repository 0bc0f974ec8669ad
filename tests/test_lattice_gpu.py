"""GPU parity of the sparse lattice product and of full MIA contractions
(through the C-ABI and the SparseMIA mirror) against the CPU oracle: keys
bit-exact, values within 1e-12 under isEqualFuzzy (FunctionUtil.h:688-698),
because only the summation order of duplicate rows differs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def _random_lattice(orc, h, w, d, n, seed):
    n = min(n, h * w * d)
    k = np.sort(orc.gen_distinct_keys(n, h * w * d, seed))
    v = orc.gen_uniform_vals(n, seed + 1)
    return k, v


CASES = [
    # (ah, aw, d, bw, nA, nB)
    (1024, 1024, 1, 1024, 10486, 10486),  # c1 shape
    (50, 40, 3, 30, 1500, 900),  # several tabs, dense-ish
    (1 << 20, 64, 1, 4096, 200000, 3000),  # tall A columns: large segments
    (7, 5, 2, 9, 30, 40),
    (1, 1, 1, 1, 1, 1),
    (3000, 1, 4, 5000, 3000, 4000),  # no inner index (outer/inter product)
    (100, 200, 1, 1, 2000, 150),  # single B column
]


# sort, hash, dense (csc_times), esc (outer_times), mtt, and the CSC column
# lookup with sort and dense.
STRATEGIES = [1, 2, 3, 4, 5, 1 | 16, 3 | 16]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("strategy", STRATEGIES)
def test_lattice_mult_matches_oracle(ctx, orc, case, strategy):
    """Every accumulator against the oracle's csc_no_accum (keys exact, values
    to 1e-12); all but the hash accumulator against the oracle's csc_times
    sparse accumulator bit for bit (each row summed in ascending B row order),
    and outer_times/mtt_times restated agree on the keys."""
    ah, aw, d, bw, n_a, n_b = case
    print("lattice", case, "strategy", strategy, flush=True)
    ak, av = _random_lattice(orc, ah, aw, d, n_a, 3)
    bk, bv = _random_lattice(orc, aw, bw, d, n_b, 5)
    ek, ev = orc.lattice_mult(ak, av, (ah, aw, d), bk, bv, (aw, bw, d))
    nprod = orc.lattice_nprod(ak, (ah, aw, d), bk, (aw, bw, d))
    assert ctx.lattice_nprod(gpu(ak), (ah, aw, d), gpu(bk), (aw, bw, d)) == nprod
    ck, cv = ctx.lattice_mult(gpu(ak), gpu(av), (ah, aw, d), gpu(bk), gpu(bv), (aw, bw, d), strategy)
    assert np.array_equal(host(ck), ek)
    assert orc.allclose_fuzzy(host(cv), ev, TOL)
    if strategy != 2:
        sk, sv = orc.lattice_mult_by("spa", ak, av, (ah, aw, d), bk, bv, (aw, bw, d))
        assert np.array_equal(sk, ek)
        assert np.array_equal(host(cv).view(np.int64), sv.view(np.int64))
    if strategy in (4, 5):
        ok, ov = orc.lattice_mult_by("outer" if strategy == 4 else "mtt", ak, av, (ah, aw, d), bk, bv, (aw, bw, d))
        assert np.array_equal(ok, ek) and orc.allclose_fuzzy(ov, ev, TOL)


@pytest.mark.parametrize("shape", [(8192, 300, 2, 400), (8193, 300, 1, 400), (50000, 64, 1, 2000),
                                   (1 << 20, 16, 1, 1000)])
def test_dense_and_mtt_limits(ctx, orc, shape):
    """The dense accumulators at and past their LDS limit (8192 rows, or 8192
    distinct rows for mtt_times), with long columns (many B entries each):
    past the limit they fall back to the sort accumulator; the bits stay the
    sparse accumulator's either way."""
    ah, aw, d, bw = shape
    rng = np.random.default_rng(ah)
    n_a = min(60000, ah * aw * d // 4)
    ak, av = _random_lattice(orc, ah, aw, d, n_a, 7)
    if ah > 20000:  # few distinct rows: mtt's ranks fit where the rows do not
        rows = rng.choice(ah, 6000, replace=False)
        sub = np.unique(rows[rng.integers(0, 6000, n_a)] + ah * rng.integers(0, aw * d, n_a))
        ak, av = sub.astype(np.int64), orc.gen_uniform_vals(len(sub), 8)
    # B: a few columns with many entries (long segments) and many short ones
    bk = np.unique(np.concatenate([rng.integers(0, aw, 3000) + aw * rng.integers(0, 4, 3000),
                                   rng.integers(0, aw * bw, 5000)]) % (aw * bw * d)).astype(np.int64)
    bv = orc.gen_uniform_vals(len(bk), 9)
    sk, sv = orc.lattice_mult_by("spa", ak, av, (ah, aw, d), bk, bv, (aw, bw, d))
    for strategy in (3, 5, 3 | 16):
        print("dense limits", shape, "strategy", strategy, flush=True)
        ck, cv = ctx.lattice_mult(gpu(ak), gpu(av), (ah, aw, d), gpu(bk), gpu(bv), (aw, bw, d), strategy)
        assert np.array_equal(host(ck), sk), strategy
        assert np.array_equal(host(cv).view(np.int64), sv.view(np.int64)), strategy


def test_lattice_mult_skewed_columns(ctx, orc):
    """B columns with many entries over long A columns exercise every segment
    kind (single entry, LDS, global sort)."""
    rng = np.random.default_rng(4)
    ah, aw, bw = 1 << 16, 256, 512
    a_cols = rng.integers(0, aw, 60000)
    a_rows = rng.integers(0, ah, 60000)
    ak = np.unique(a_rows + ah * a_cols).astype(np.int64)
    av = rng.uniform(-1, 1, len(ak))
    # B: column 0 holds every k, columns 1..40 a handful, the rest one entry
    ks, cs = [], []
    ks += list(range(aw)); cs += [0] * aw
    for c in range(1, 41):
        sel = rng.choice(aw, 6, replace=False)
        ks += list(sel); cs += [c] * 6
    for c in range(41, bw):
        ks.append(int(rng.integers(0, aw))); cs.append(c)
    bk = np.unique(np.array(ks) + aw * np.array(cs)).astype(np.int64)
    bv = rng.uniform(-1, 1, len(bk))
    ek, ev = orc.lattice_mult(ak, av, (ah, aw, 1), bk, bv, (aw, bw, 1))
    for strategy in (1, 2):
        ck, cv = ctx.lattice_mult(gpu(ak), gpu(av), (ah, aw, 1), gpu(bk), gpu(bv), (aw, bw, 1), strategy)
        assert np.array_equal(host(ck), ek)
        assert orc.allclose_fuzzy(host(cv), ev, TOL)


def test_lattice_mult_sort_strategy_is_deterministic(ctx, orc):
    ak, av = _random_lattice(orc, 4096, 512, 1, 300000, 8)
    bk, bv = _random_lattice(orc, 512, 2048, 1, 20000, 9)
    outs = [ctx.lattice_mult(gpu(ak), gpu(av), (4096, 512, 1), gpu(bk), gpu(bv), (512, 2048, 1), 1) for _ in range(3)]
    for ck, cv in outs[1:]:
        assert np.array_equal(host(ck), host(outs[0][0]))
        assert np.array_equal(host(cv).view(np.int64), host(outs[0][1]).view(np.int64))


@pytest.mark.parametrize("shape", [(1 << 20, 64, 1, 4096, 200000, 3000), (4096 * 4096, 512, 1, 4096 * 64, 858993, 13422)])
def test_lattice_mult_row_ranks(ctx, orc, shape):
    """Long output columns sorted by row rank instead of row: the same keys
    and the same value bits as with rows (same sort order, same summation
    order), for both accumulators, and the oracle's result."""
    ah, aw, d, bw, n_a, n_b = shape
    print("row ranks", shape, flush=True)
    ak, av = _random_lattice(orc, ah, aw, d, n_a, 21)
    bk, bv = _random_lattice(orc, aw, bw, d, n_b, 22)
    ek, ev = orc.lattice_mult(ak, av, (ah, aw, d), bk, bv, (aw, bw, d))
    ctx.set_param("spgemm_pieces", 0)  # every long column through the global sort
    try:
        for strategy in (1, 2):
            outs = []
            for ranks in (1, 0):
                ctx.set_param("row_ranks", ranks)
                outs.append(ctx.lattice_mult(gpu(ak), gpu(av), (ah, aw, d), gpu(bk), gpu(bv), (aw, bw, d), strategy))
            assert np.array_equal(host(outs[0][0]), host(outs[1][0]))
            if strategy == 1:  # the sort accumulator is deterministic
                assert np.array_equal(host(outs[0][1]).view(np.int64), host(outs[1][1]).view(np.int64))
            assert np.array_equal(host(outs[0][0]), ek)
            assert orc.allclose_fuzzy(host(outs[0][1]), ev, TOL)
    finally:
        ctx.set_param("row_ranks", 1)
        ctx.set_param("spgemm_pieces", 1)


def _piece_case(kind, rng):
    """(ah, aw, bw, ak, av, bk, bv) of one tab, built to steer the long
    output columns through a given branch of the row-piece path."""
    ah, aw, bw = 1 << 20, 96, 8
    rows = []
    for col in range(aw):
        if kind == "uniform":  # pieces, every column in one piece range
            r = rng.choice(ah, 3000, replace=False)
        elif kind == "gap":  # rows in two bands: pieces over the gap are empty
            r = np.concatenate([rng.choice(30000, 9000, replace=False),
                                40000 + rng.choice(30000, 9000, replace=False)])
        elif kind == "cluster":  # one far row: the first piece overflows and splits itself
            r = np.concatenate([rng.choice(2000, 1500, replace=False), [ah - 1]])
        else:  # "equal": every column the same rows -> runs of equal rows across all B entries
            r = np.arange(0, ah, ah // 2500)
        rows.append(np.unique(r) + ah * col)
    ak = np.concatenate(rows).astype(np.int64)
    av = rng.uniform(-1, 1, len(ak))
    ks, cs = [], []
    for c, cnt in enumerate([2, 5, 17, 64, 65, 96, 1, 40]):  # B entries per output column
        ks += list(np.sort(rng.choice(aw, cnt, replace=False)))
        cs += [c] * cnt
    bk = (np.array(ks) + aw * np.array(cs)).astype(np.int64)
    order = np.argsort(bk)
    bk = bk[order]
    bv = rng.uniform(-1, 1, len(bk))
    return ah, aw, bw, ak, av, bk, bv


@pytest.mark.parametrize("kind", ["uniform", "gap", "cluster", "equal"])
def test_lattice_mult_row_pieces(ctx, orc, kind):
    """Long output columns in row pieces, with 2..64 B entries per column
    (pieces), 65 and 96 (global sort), uniform rows, two bands with empty
    pieces between, a cluster that overflows a piece (split into row
    sub-ranges inside its workgroup), and
    identical rows in every column (runs across all B entries): keys exact
    and value bits equal to the oracle's sparse accumulator and to the
    global sort's, for every accumulator that uses the pieces.  The pieces
    run twice: writing straight into C (output arrays of nprod entries) and
    staged through the workspace (arrays of exactly nnz(C) entries)."""
    rng = np.random.default_rng(["uniform", "gap", "cluster", "equal"].index(kind) + 40)
    ah, aw, bw, ak, av, bk, bv = _piece_case(kind, rng)
    print("row pieces", kind, len(ak), len(bk), flush=True)
    sk, sv = orc.lattice_mult_by("spa", ak, av, (ah, aw, 1), bk, bv, (aw, bw, 1))
    try:
        for strategy in (1, 2, 17):
            for mode in ("direct", "staged", "sort"):
                ctx.set_param("spgemm_pieces", int(mode != "sort"))
                if mode == "staged":
                    import torch

                    ck = torch.empty(len(sk), dtype=torch.int64, device="cuda")
                    cv = torch.empty(len(sk), dtype=torch.float64, device="cuda")
                    n = ctx.lattice_mult_into(gpu(ak), gpu(av), (ah, aw, 1), gpu(bk), gpu(bv), (aw, bw, 1),
                                              ck, cv, strategy)
                    assert n == len(sk), (strategy, mode)
                else:
                    ck, cv = ctx.lattice_mult(gpu(ak), gpu(av), (ah, aw, 1), gpu(bk), gpu(bv), (aw, bw, 1), strategy)
                assert np.array_equal(host(ck), sk), (strategy, mode)
                if strategy != 2:
                    assert np.array_equal(host(cv).view(np.int64), sv.view(np.int64)), (strategy, mode)
                else:
                    assert orc.allclose_fuzzy(host(cv), sv, TOL)
    finally:
        ctx.set_param("spgemm_pieces", 1)


@pytest.mark.parametrize("others", ["none", "short", "long", "both"])
def test_lattice_mult_single_entry_columns(ctx, orc, others):
    """Output columns of one B entry (a scaled copy of one A column) long
    enough on average for the XCD-local copy kernel: A columns of 0 .. 3,000
    entries, alone or beside short columns (LDS accumulators, staged), long
    ones (row pieces) or both.  Keys and value bits equal the oracle's
    sparse accumulator with and without the copy kernel, into arrays of
    nprod entries (launches queued ahead, pieces direct), of exactly nnz(C)
    entries (pieces staged) and allocated."""
    import torch

    rng = np.random.default_rng(["none", "short", "long", "both"].index(others) + 120)
    ah = 8192
    lens = [0, 1, 2, 5, 63, 64, 65, 127, 128, 129, 511, 512, 513, 1000, 3000, 2500] * 4
    aw = len(lens)
    rows = [np.sort(rng.choice(ah, n, replace=False)).astype(np.int64) + ah * c for c, n in enumerate(lens)]
    ak = np.concatenate(rows).astype(np.int64)
    av = rng.uniform(-1, 1, len(ak))
    av[::41] = -0.0
    counts = [1] * 300
    if others in ("short", "both"):
        counts += [2, 3] * 10
    if others in ("long", "both"):
        counts += [6, 12, 30] * 4
    rng.shuffle(counts)
    ks, cs = [], []
    for c, cnt in enumerate(counts):
        ks += list(np.sort(rng.choice(aw, cnt, replace=False)))
        cs += [c] * cnt
    bw = len(counts)
    bk = (np.array(ks) + aw * np.array(cs)).astype(np.int64)
    order = np.argsort(bk)
    bk = bk[order]
    bv = rng.uniform(-1, 1, len(bk))
    bv[::7] = -0.0
    sk, sv = orc.lattice_mult_by("spa", ak, av, (ah, aw, 1), bk, bv, (aw, bw, 1))
    nprod = int(orc.lattice_nprod(ak, (ah, aw, 1), bk, (aw, bw, 1)))
    print("single-entry columns", others, len(ak), len(bk), nprod, len(sk), flush=True)
    try:
        for single_copy in (1, 0):
            ctx.set_param("spgemm_single_copy", single_copy)
            for strategy in (1, 17):
                for cap in (nprod, len(sk), None):
                    if cap is None:
                        ck, cv = ctx.lattice_mult(gpu(ak), gpu(av), (ah, aw, 1), gpu(bk), gpu(bv), (aw, bw, 1),
                                                  strategy)
                    else:
                        ck = torch.empty(cap, dtype=torch.int64, device="cuda")
                        cv = torch.empty(cap, dtype=torch.float64, device="cuda")
                        n = ctx.lattice_mult_into(gpu(ak), gpu(av), (ah, aw, 1), gpu(bk), gpu(bv), (aw, bw, 1),
                                                  ck, cv, strategy)
                        assert n == len(sk)
                        ck, cv = ck[:n], cv[:n]
                    tag = (single_copy, strategy, cap)
                    assert np.array_equal(host(ck), sk), tag
                    assert np.array_equal(host(cv).view(np.int64), sv.view(np.int64)), tag
    finally:
        ctx.set_param("spgemm_single_copy", 1)


@pytest.mark.parametrize("extra_short", [False, True])
def test_lattice_mult_all_long_columns_fall_back(ctx, orc, extra_short):
    """Every long output column has more than 64 B entries, so the row pieces
    take none of them and all go to the global sort (the contraction finds
    this at the end and runs again waiting for the lists); with and without
    a single-entry column beside them.  Keys and value bits equal the
    oracle's sparse accumulator, into arrays of nprod entries (output
    launches queued ahead) and of exactly nnz(C) entries."""
    import torch

    rng = np.random.default_rng(91 + int(extra_short))
    ah, aw = 300, 200
    rows = [np.sort(rng.choice(ah, int(rng.integers(20, 120)), replace=False)) + ah * c for c in range(aw)]
    ak = np.concatenate(rows).astype(np.int64)
    av = rng.uniform(-1, 1, len(ak))
    counts = [100, 130, 65] + ([1] if extra_short else [])
    ks, cs = [], []
    for c, cnt in enumerate(counts):
        ks += list(np.sort(rng.choice(aw, cnt, replace=False)))
        cs += [c] * cnt
    bk = (np.array(ks) + aw * np.array(cs)).astype(np.int64)
    bv = rng.uniform(-1, 1, len(bk))
    bw = len(counts)
    sk, sv = orc.lattice_mult_by("spa", ak, av, (ah, aw, 1), bk, bv, (aw, bw, 1))
    print("all-fallback columns", extra_short, len(ak), len(bk), flush=True)
    for strategy in (1, 17):
        ck, cv = ctx.lattice_mult(gpu(ak), gpu(av), (ah, aw, 1), gpu(bk), gpu(bv), (aw, bw, 1), strategy)
        assert np.array_equal(host(ck), sk), strategy
        assert np.array_equal(host(cv).view(np.int64), sv.view(np.int64)), strategy
        ck = torch.empty(len(sk), dtype=torch.int64, device="cuda")
        cv = torch.empty(len(sk), dtype=torch.float64, device="cuda")
        n = ctx.lattice_mult_into(gpu(ak), gpu(av), (ah, aw, 1), gpu(bk), gpu(bv), (aw, bw, 1), ck, cv, strategy)
        assert n == len(sk)
        assert np.array_equal(host(ck), sk) and np.array_equal(host(cv).view(np.int64), sv.view(np.int64))


def test_lattice_mult_empty_and_errors(ctx, orc):
    import libnt_amd
    import torch

    e = torch.empty(0, dtype=torch.int64, device="cuda")
    ev = torch.empty(0, dtype=torch.float64, device="cuda")
    ak, av = _random_lattice(orc, 10, 10, 1, 20, 1)
    ck, cv = ctx.lattice_mult(e, ev, (10, 10, 1), gpu(ak), gpu(av), (10, 10, 1))
    assert ck.shape[0] == 0
    with pytest.raises(libnt_amd.LntError) as err:
        ctx.lattice_mult(gpu(ak), gpu(av), (10, 10, 1), gpu(ak), gpu(av), (9, 10, 1))
    assert err.value.status == 7


def test_lattice_mult_host_memory(ctx, orc):
    ak, av = _random_lattice(orc, 300, 200, 2, 5000, 11)
    bk, bv = _random_lattice(orc, 200, 100, 2, 3000, 12)
    ek, ev = orc.lattice_mult(ak, av, (300, 200, 2), bk, bv, (200, 100, 2))
    ck, cv = ctx.lattice_mult(ak, av, (300, 200, 2), bk, bv, (200, 100, 2))
    assert np.array_equal(ck, ek) and orc.allclose_fuzzy(cv, ev, TOL)


# ------------------------------------------------ full MIA contractions
def _c1_operand(orc, seed):
    keys = orc.gen_distinct_keys(10486, 1 << 20, seed)
    vals = orc.gen_uniform_vals(10486, seed + 1000)
    return keys, vals


def test_c1_contraction(orc):
    """C(i,j,m,n) = A(i,j,k,l) * B(k,l,m,n), dims 32, 1% density, unsorted
    operands in the default order."""
    from libnt_amd.sparse import SparseMIA

    ak, av = _c1_operand(orc, 1)
    bk, bv = _c1_operand(orc, 2)
    dims = [32, 32, 32, 32]
    ek, ev, edims = orc.contract(ak, av, dims, [0, 1, 2, 3], False, "ijkl", bk, bv, dims, [0, 1, 2, 3], False, "klmn")
    A = SparseMIA(dims, ak, av)
    B = SparseMIA(dims, bk, bv)
    C = (A("ijkl") * B("klmn")).mia
    assert list(C.dims) == list(edims)
    assert np.array_equal(host(C.keys), ek)
    assert orc.allclose_fuzzy(host(C.vals), ev, TOL)


@pytest.mark.parametrize(
    "la,lb,dims_a,dims_b",
    [
        ("kij", "jlk", [40, 50, 60], [60, 70, 40]),
        ("iJk", "kJm", [50, 30, 40], [40, 30, 60]),
        ("ij", "kl", [30, 40], [50, 20]),
        ("IJK", "KIJ", [30, 20, 10], [10, 30, 20]),
        ("aIbc", "cbdI", [9, 8, 7, 6], [6, 7, 5, 8]),
    ],
)
def test_contractions_match_oracle(orc, la, lb, dims_a, dims_b):
    from libnt_amd.sparse import SparseMIA

    na = max(1, int(0.05 * np.prod(dims_a)))
    nb = max(1, int(0.05 * np.prod(dims_b)))
    ak = orc.gen_distinct_keys(na, int(np.prod(dims_a)), 5)
    av = orc.gen_uniform_vals(na, 6)
    bk = orc.gen_distinct_keys(nb, int(np.prod(dims_b)), 7)
    bv = orc.gen_uniform_vals(nb, 8)
    ident_a, ident_b = list(range(len(dims_a))), list(range(len(dims_b)))
    ek, ev, edims = orc.contract(ak, av, dims_a, ident_a, False, la, bk, bv, dims_b, ident_b, False, lb)
    C = (SparseMIA(dims_a, ak, av)(la) * SparseMIA(dims_b, bk, bv)(lb)).mia
    assert list(C.dims) == list(edims)
    assert np.array_equal(host(C.keys), ek)
    assert orc.allclose_fuzzy(host(C.vals), ev, TOL)


def test_front_end_merge_rules(orc):
    """C(i,j,k,l) = A(i,j,k,l) + B(k,l,i,j) against dense arithmetic, for every
    combination of sorted flags (SparseMIABase.h:1404-1455)."""
    from libnt_amd.sparse import SparseMIA

    dims_a = [6, 5, 4, 3]
    dims_b = [4, 3, 6, 5]
    ak = orc.gen_distinct_keys(150, 360, 1)
    av = orc.gen_uniform_vals(150, 2)
    bk = orc.gen_distinct_keys(200, 360, 3)
    bv = orc.gen_uniform_vals(200, 4)
    for a_sorted in (False, True):
        for b_sorted in (False, True):
            A = SparseMIA(dims_a, ak, av)
            B = SparseMIA(dims_b, bk, bv)
            if a_sorted:
                A.sort([2, 0, 1, 3])
            if b_sorted:
                B.sort([3, 1, 0, 2])
            Ad, Bd = A.to_dense(), B.to_dense()
            Bt = Bd.transpose(2, 3, 0, 1)  # B(k,l,i,j) as (i,j,k,l)
            for op, fn in (("+", np.add), ("-", np.subtract)):
                C = A("ijkl") + B("klij") if op == "+" else A("ijkl") - B("klij")
                assert C.dims == tuple(dims_a)
                D = C.to_dense()
                assert np.array_equal(D, fn(Ad, Bt))
                assert C.nnz == int(np.count_nonzero((Ad != 0) | (Bt != 0)))


def test_front_end_hadamard(orc):
    from libnt_amd.sparse import SparseMIA

    dims = [16, 16, 16]
    ak = orc.gen_distinct_keys(1500, 4096, 1)
    av = orc.gen_uniform_vals(1500, 2)
    bk = orc.gen_distinct_keys(1500, 4096, 3)
    bv = orc.gen_uniform_vals(1500, 4)
    A, B = SparseMIA(dims, ak, av), SparseMIA(dims, bk, bv)
    Ad, Bd = A.to_dense(), B.to_dense()
    C = (A("IJK") * B("KIJ")).mia
    D = C.to_dense()
    expect = Ad * Bd.transpose(1, 2, 0)
    assert np.array_equal(D, expect)


@pytest.mark.parametrize("strategy", [1, 2])
def test_c3_skewed_contraction_full_size(orc, strategy):
    """A(i,j,k) * B(k,l,m): A 4096x4096x512, B 512x4096x64, density 1e-4, i and
    l Zipf(1.1), values N(0,1) -- the full-size skewed configuration."""
    import torch

    from libnt_amd.sparse import SparseMIA

    dims_a, dims_b = [4096, 4096, 512], [512, 4096, 64]
    ak, av = orc.gen_coords(858993, dims_a, [1, 0, 0], 1.1, 1, 31)
    bk, bv = orc.gen_coords(13422, dims_b, [0, 1, 0], 1.1, 1, 32)
    ek, ev, edims = orc.contract(ak, av, dims_a, [0, 1, 2], False, "ijk", bk, bv, dims_b, [0, 1, 2], False, "klm")
    A, B = SparseMIA(dims_a, ak, av), SparseMIA(dims_b, bk, bv)
    A.sort([0, 1, 2])
    B.sort([0, 1, 2])
    ck, cv = A.ctx.lattice_mult(A.keys, A.vals, (4096 * 4096, 512, 1), B.keys, B.vals, (512, 4096 * 64, 1), strategy)
    assert np.array_equal(host(ck), ek)
    assert orc.allclose_fuzzy(host(cv), ev, TOL)
