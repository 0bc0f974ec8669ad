"""GPU parity of batched contractions (include/lnt_batch.h) through the C-ABI:
every instance against the CPU oracle's full contraction (keys bit-exact,
values within 1e-12 under isEqualFuzzy, FunctionUtil.h:688-698) and against
the single-contraction GPU path (bit-identical: both sum equal rows in
ascending B-row order); the recorded hipGraph against the eager run; and the
full-size mix through size-independent properties (per-instance sortedness,
nnz bounds, the linearity checksum sum(C) = sum_k colsum_A(k) * rowsum_B(k),
and a sample of its instances against the oracle entry by entry)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def host(t):
    return t.cpu().numpy()


def _oracle_operands(orc, insts, i):
    from libnt_amd import workloads as W

    x = insts[i]
    ak = orc.gen_distinct_keys(x.nnz_a, x.dimensionality_a, W.KEY_SEED_A + i)
    av = orc.gen_uniform_vals(x.nnz_a, W.VAL_SEED_A + i)
    bk = orc.gen_distinct_keys(x.nnz_b, x.dimensionality_b, W.KEY_SEED_B + i)
    bv = orc.gen_uniform_vals(x.nnz_b, W.VAL_SEED_B + i)
    return ak, av, bk, bv


def _small_mix(count=120, seed=9):
    from libnt_amd import workloads as W

    return W.batched_contractions(count, seed=seed, nnz_lo=10, nnz_hi=3000, dim_lo=3, dim_hi=24)


def _run(ctx, insts):
    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    ops = W.generate_operands(ctx, insts)
    b = Batch(ctx, insts)
    ck, cv, off = b.contract(*ops)
    return b, ops, ck, cv, off


def test_generated_operands_match_oracle(ctx, orc):
    insts = _small_mix(40, seed=3)
    from libnt_amd import workloads as W
    from libnt_amd.batch import packed_offsets

    ak, av, bk, bv = W.generate_operands(ctx, insts)
    a_off = packed_offsets([x.nnz_a for x in insts])
    b_off = packed_offsets([x.nnz_b for x in insts])
    for i in range(len(insts)):
        eak, eav, ebk, ebv = _oracle_operands(orc, insts, i)
        assert np.array_equal(host(ak[a_off[i]:a_off[i + 1]]), eak)
        assert np.array_equal(host(av[a_off[i]:a_off[i + 1]]).view(np.int64), eav.view(np.int64))
        assert np.array_equal(host(bk[b_off[i]:b_off[i + 1]]), ebk)
        assert np.array_equal(host(bv[b_off[i]:b_off[i + 1]]).view(np.int64), ebv.view(np.int64))


@pytest.mark.parametrize("seed", [9, 10])
def test_batch_matches_oracle(ctx, orc, seed):
    insts = _small_mix(120, seed=seed)
    print("batch vs oracle, seed", seed, flush=True)
    b, ops, ck, cv, off = _run(ctx, insts)
    ck, cv = host(ck), host(cv)
    for i, x in enumerate(insts):
        ak, av, bk, bv = _oracle_operands(orc, insts, i)
        ident_a, ident_b = list(range(len(x.dims_a))), list(range(len(x.dims_b)))
        ek, ev, edims = orc.contract(ak, av, x.dims_a, ident_a, False, x.labels_a,
                                     bk, bv, x.dims_b, ident_b, False, x.labels_b)
        assert list(edims) == x.dims_c
        got_k, got_v = ck[off[i]:off[i + 1]], cv[off[i]:off[i + 1]]
        assert np.array_equal(got_k, ek), "instance %d keys" % i
        assert orc.allclose_fuzzy(got_v, ev, TOL), "instance %d values" % i


def test_batch_nprod_matches_oracle(ctx, orc):
    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    insts = _small_mix(60, seed=12)
    ak, av, bk, bv = W.generate_operands(ctx, insts)
    got = Batch(ctx, insts).nprod(ak, bk)
    for i, x in enumerate(insts):
        a, _, bb, _ = _oracle_operands(orc, insts, i)
        ident_a, ident_b = list(range(len(x.dims_a))), list(range(len(x.dims_b)))
        a2, _ = orc.sort_seq(a, np.zeros(len(a)), x.dims_a, ident_a, x.seq_a, False)
        b2, _ = orc.sort_seq(bb, np.zeros(len(bb)), x.dims_b, ident_b, x.seq_b, False)
        assert got[i] == orc.lattice_nprod(a2, x.lat_a, b2, x.lat_b), i


def _single_path(ctx, x, ak, av, bk, bv):
    """One instance through the single-contraction path (SparseMIA mirror)."""
    from libnt_amd.sparse import SparseMIA

    A = SparseMIA(x.dims_a, ak.clone(), av.clone(), ctx=ctx)
    B = SparseMIA(x.dims_b, bk.clone(), bv.clone(), ctx=ctx)
    C = (A(x.labels_a) * B(x.labels_b)).mia
    assert list(C.dims) == x.dims_c
    return C.keys, C.vals


def test_batch_matches_single_contraction_path(ctx):
    """A slice of the full-size mix (every size class of columns) equals the
    single-contraction path bit for bit."""
    from libnt_amd import workloads as W
    from libnt_amd.batch import packed_offsets

    insts = W.batched_contractions(300, seed=4)
    b, (ak, av, bk, bv), ck, cv, off = _run(ctx, insts)
    a_off = packed_offsets([x.nnz_a for x in insts])
    b_off = packed_offsets([x.nnz_b for x in insts])
    for i, x in enumerate(insts):
        sk, sv = _single_path(ctx, x, ak[a_off[i]:a_off[i + 1]], av[a_off[i]:a_off[i + 1]],
                              bk[b_off[i]:b_off[i + 1]], bv[b_off[i]:b_off[i + 1]])
        import torch

        assert torch.equal(ck[off[i]:off[i + 1]], sk), "instance %d keys" % i
        assert torch.equal(cv[off[i]:off[i + 1]].view(torch.int64), sv.view(torch.int64)), "instance %d values" % i


def test_batch_tuning_does_not_change_results(ctx):
    import torch

    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    insts = W.batched_contractions(120, seed=21)
    ops = W.generate_operands(ctx, insts)
    b = Batch(ctx, insts)
    ref = b.contract(*ops)
    for params in ({"small_max": 2, "item_products": 64, "window": 64, "blocks_per_cu": 1},
                   {"small_max": 1024, "item_products": 1024, "window": 1024, "blocks_per_cu": 8},
                   {"small_max": 300, "item_products": 200, "window": 100, "blocks_per_cu": 2}):
        for k, v in params.items():
            b.set_param(k, v)
        ck, cv, off = b.contract(*ops)
        assert np.array_equal(off, ref[2])
        assert torch.equal(ck, ref[0]), params
        # A window change moves sums of one row between the sort order and
        # the tree order only for rows met by > 2048 products: none here.
        assert torch.equal(cv.view(torch.int64), ref[1].view(torch.int64)), params


def test_batch_row_pieces_match_windows(ctx):
    """Large segments of at most 64 B entries go through row pieces; with no
    piece slots they take the windows, and with a few slots some take each:
    keys, offsets and value bits equal in every case, eager and replayed."""
    import torch

    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    insts = W.batched_contractions(160, seed=23)
    ops = W.generate_operands(ctx, insts)
    b = Batch(ctx, insts)
    ref = b.contract(*ops)  # row pieces
    for slots in (0, 5, 300, -1):
        b.set_param("piece_slots", slots)
        ck, cv, off = b.contract(*ops)
        assert np.array_equal(off, ref[2]), slots
        assert torch.equal(ck, ref[0]), slots
        assert torch.equal(cv.view(torch.int64), ref[1].view(torch.int64)), slots
    b.set_param("piece_slots", 7)
    n = b.capture(*ops)
    assert n == ref[0].shape[0]
    for rep in range(2):
        gk, gv, goff = b.replay()
        assert np.array_equal(goff, ref[2])
        assert torch.equal(gk, ref[0]) and torch.equal(gv.view(torch.int64), ref[1].view(torch.int64)), rep


@pytest.mark.parametrize("kind", ["uniform", "gap", "cluster", "equal"])
def test_batch_row_pieces_shapes(ctx, orc, kind):
    """Long output columns built to steer the batch's row pieces through each
    branch (2..64 B entries: pieces; 65 and 96: windows; empty pieces over a
    gap; a cluster that overflows a piece and splits into row sub-ranges;
    equal rows in every column), as a batch of three instances: keys and
    value bits equal to the oracle's sparse accumulator."""
    import importlib.util
    import os

    import torch

    from libnt_amd.batch import Batch, Instance

    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("lattice_cases", os.path.join(here, "test_lattice_gpu.py"))
    cases = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cases)
    rng = np.random.default_rng(["uniform", "gap", "cluster", "equal"].index(kind) + 60)
    insts, parts, expect = [], [[], [], [], []], []
    for _ in range(3):
        ah, aw, bw, ak, av, bk, bv = cases._piece_case(kind, rng)
        insts.append(Instance([ah, aw], "ca", [aw, bw], "ag", len(ak), len(bk)))
        for lst, x in zip(parts, (ak, av, bk, bv)):
            lst.append(x)
        expect.append(orc.lattice_mult_by("spa", ak, av, (ah, aw, 1), bk, bv, (aw, bw, 1)))
    print("batch row pieces", kind, [len(x) for x in parts[0]], flush=True)
    ops = [torch.from_numpy(np.concatenate(lst)).cuda() for lst in parts]
    b = Batch(ctx, insts)
    for slots in (-1, 0):
        b.set_param("piece_slots", slots)
        ck, cv, off = b.contract(*ops)
        for i, (ek, ev) in enumerate(expect):
            k = host(ck[off[i]:off[i + 1]])
            v = host(cv[off[i]:off[i + 1]])
            assert np.array_equal(k, ek), (slots, i)
            assert np.array_equal(v.view(np.int64), ev.view(np.int64)), (slots, i)


def test_batch_single_entry_columns(ctx, orc):
    """Output columns of one B entry (a scaled copy of one A column), spread
    over the lanes of a wave 64 columns at a time: A columns of 0 .. 3,000
    entries (around the wave width and the 512-product chunk), runs of
    one-entry columns broken by columns of two and three entries, more
    instances than XCD ranges and a first instance with none: keys and value
    bits equal to the oracle's sparse accumulator, eager and replayed."""
    import torch

    from libnt_amd.batch import Batch, Instance

    rng = np.random.default_rng(31)
    lens = [0, 1, 2, 5, 63, 64, 65, 127, 128, 129, 511, 512, 513, 600, 1000, 3000]
    insts, parts, expect = [], [[], [], [], []], []
    for inst in range(11):
        h, aw = 4096, len(lens)
        bw = 1 if inst == 0 else int(rng.integers(50, 400))
        rows = [np.sort(rng.choice(h, n, replace=False)).astype(np.int64) + h * c for c, n in enumerate(lens)]
        ak = np.concatenate(rows).astype(np.int64)
        av = rng.uniform(-1, 1, len(ak))
        av[::37] = -0.0
        ks, cs = [], []
        for c in range(bw):
            cnt = 2 if inst == 0 else (1 if rng.random() < 0.85 else int(rng.integers(2, 4)))
            ks += list(np.sort(rng.choice(aw, cnt, replace=False)))
            cs += [c] * cnt
        bk = (np.array(ks) + aw * np.array(cs)).astype(np.int64)
        order = np.argsort(bk)
        bk = bk[order]
        bv = rng.uniform(-1, 1, len(bk))
        bv[::11] = -0.0
        insts.append(Instance([h, aw], "ca", [aw, bw], "ag", len(ak), len(bk)))
        for lst, x in zip(parts, (ak, av, bk, bv)):
            lst.append(x)
        expect.append(orc.lattice_mult_by("spa", ak, av, (h, aw, 1), bk, bv, (aw, bw, 1)))
    print("batch single-entry columns", [len(x) for x in parts[2]], flush=True)
    ops = [torch.from_numpy(np.concatenate(lst)).cuda() for lst in parts]
    b = Batch(ctx, insts)

    def check_all(ck, cv, off):
        for i, (ek, ev) in enumerate(expect):
            k = host(ck[off[i]:off[i + 1]])
            v = host(cv[off[i]:off[i + 1]])
            assert np.array_equal(k, ek), i
            assert np.array_equal(v.view(np.int64), ev.view(np.int64)), i

    check_all(*b.contract(*ops))
    b.capture(*ops)
    for _ in range(2):
        check_all(*b.replay())


@pytest.mark.parametrize("seed", [21, 23])
def test_batch_accumulator_choices_do_not_change_results(ctx, seed):
    """The per-instance accumulator choice (lnt_batch_set_accumulators):
    every instance dense, a seeded random mix of choices and the smallest
    dense segments allowed give keys, offsets and value bits equal to the
    batch's own merge/sort paths, eager and in a recorded graph whose
    replays follow later choices."""
    import torch

    from libnt_amd import dispatch as D
    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    insts = W.batched_contractions(160, seed=seed)
    ops = W.generate_operands(ctx, insts)
    b = Batch(ctx, insts)
    ref = b.contract(*ops)
    assert b.get_param("dense_segments") == 0
    rng = np.random.default_rng(seed)
    picks = {
        "dense": np.full(len(insts), D.ACC_DENSE),
        "mixed": rng.integers(0, D.NUM_STRATEGIES // 4, len(insts)),
    }
    for name, acc in picks.items():
        for dense_min in (256, 16):
            print("accumulators", seed, name, dense_min, flush=True)
            b.set_accumulators(acc)
            b.set_param("dense_min", dense_min)
            ck, cv, off = b.contract(*ops)
            assert b.get_param("dense_segments") > 0, (name, dense_min)
            assert np.array_equal(off, ref[2]), (name, dense_min)
            assert torch.equal(ck, ref[0]), (name, dense_min)
            assert torch.equal(cv.view(torch.int64), ref[1].view(torch.int64)), (name, dense_min)
    b.set_accumulators(picks["dense"])
    n = b.capture(*ops)
    assert n == ref[0].shape[0]
    for acc in (picks["dense"], picks["mixed"], np.full(len(insts), D.ACC_SORT)):
        b.set_accumulators(acc)
        gk, gv, goff = b.replay()
        assert np.array_equal(goff, ref[2])
        assert torch.equal(gk, ref[0]) and torch.equal(gv.view(torch.int64), ref[1].view(torch.int64))
    assert b.get_param("dense_segments") == 0  # the last replay ran with no dense instance


def _dense_case(h, rng):
    """(aw, bw, ak, av, bk, bv) of one tab of height h: A columns empty, short,
    long (one B entry's products span several chunks of the dense kernel),
    full and identical; B output columns of 1 .. 299 entries (more than one
    table of entries per chunk)."""
    aw, bw = 300, 7
    rows = []
    for col in range(aw):
        kind = col % 6
        if kind == 0:
            r = np.zeros(0, dtype=np.int64)
        elif kind == 1:
            r = rng.choice(h, min(h, int(rng.integers(1, 5))), replace=False)
        elif kind == 2:
            r = rng.choice(h, min(h, 2500), replace=False)
        elif kind == 3:
            r = np.arange(h)
        elif kind == 4:
            r = np.arange(0, h, 3)
        else:
            r = rng.choice(h, max(1, h // 10), replace=False)
        rows.append(np.sort(r).astype(np.int64) + h * col)
    ak = np.concatenate(rows).astype(np.int64)
    av = rng.uniform(-1, 1, len(ak))
    av[::97] = -0.0  # signed zeros: a row's first product is stored, not added to 0
    ks, cs = [], []
    for c, cnt in enumerate([2, 17, 64, 65, 130, 299, 1]):
        ks += list(np.sort(rng.choice(aw, cnt, replace=False)))
        cs += [c] * cnt
    bk = (np.array(ks) + aw * np.array(cs)).astype(np.int64)
    order = np.argsort(bk)
    bk = bk[order]
    bv = rng.uniform(-1, 1, len(bk))
    bv[::13] = -0.0
    return aw, bw, ak, av, bk, bv


def test_batch_dense_accumulator_shapes(ctx, orc):
    """The dense accumulator on columns built for its branches (see
    _dense_case) at heights 33, 8191 and 8192 (the largest it takes) and 8193
    (too tall: the merge/sort paths), every instance dense and the smallest
    dense segments allowed: keys and value bits equal to the oracle's sparse
    accumulator (csc_times, ascending B-row order), eager and replayed."""
    import torch

    from libnt_amd import dispatch as D
    from libnt_amd.batch import Batch, Instance

    rng = np.random.default_rng(77)
    insts, parts, expect = [], [[], [], [], []], []
    for h in (33, 8191, 8192, 8193):
        aw, bw, ak, av, bk, bv = _dense_case(h, rng)
        insts.append(Instance([h, aw], "ca", [aw, bw], "ag", len(ak), len(bk)))
        for lst, x in zip(parts, (ak, av, bk, bv)):
            lst.append(x)
        expect.append(orc.lattice_mult_by("spa", ak, av, (h, aw, 1), bk, bv, (aw, bw, 1)))
    print("batch dense shapes", [len(x) for x in parts[0]], [len(x) for x in parts[2]], flush=True)
    ops = [torch.from_numpy(np.concatenate(lst)).cuda() for lst in parts]
    b = Batch(ctx, insts)
    b.set_accumulators(np.full(len(insts), D.ACC_DENSE))
    b.set_param("dense_min", 16)

    def check_all(ck, cv, off):
        for i, (ek, ev) in enumerate(expect):
            k = host(ck[off[i]:off[i + 1]])
            v = host(cv[off[i]:off[i + 1]])
            assert np.array_equal(k, ek), i
            assert np.array_equal(v.view(np.int64), ev.view(np.int64)), i

    # Expected dense columns: height <= 8192, two or more entries, at least
    # dense_min products and 8 per entry.
    want = 0
    for x, ak, bk in zip(insts, parts[0], parts[2]):
        h, aw = x.dims_a
        if h > 8192:
            continue
        alen = np.bincount(ak // h, minlength=aw)
        for j in np.unique(bk // aw):
            ks = bk[bk // aw == j] % aw
            npr = int(alen[ks].sum())
            want += len(ks) >= 2 and npr >= 16 and npr >= 8 * len(ks)
    assert want >= 12
    check_all(*b.contract(*ops))
    assert b.get_param("dense_segments") == want
    b.capture(*ops)
    for _ in range(2):
        check_all(*b.replay())


def test_batch_graph_replay_matches_eager(ctx):
    import torch

    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    insts = W.batched_contractions(200, seed=5)
    ak, av, bk, bv = W.generate_operands(ctx, insts)
    b = Batch(ctx, insts)
    ek, ev, eoff = b.contract(ak, av, bk, bv)
    n = b.capture(ak, av, bk, bv)
    assert n == ek.shape[0]
    for rep in range(3):
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        gk, gv, goff = b.replay()
        assert np.array_equal(goff, eoff)
        assert torch.equal(gk, ek) and torch.equal(gv.view(torch.int64), ev.view(torch.int64)), rep
        del junk
    # New values in the same buffers: the graph reads them on the next replay.
    av.mul_(-2.0)
    bv.add_(0.5)
    ek2, ev2, _ = b.contract(ak, av, bk, bv)
    gk, gv, _ = b.replay()
    assert torch.equal(gk, ek2) and torch.equal(gv.view(torch.int64), ev2.view(torch.int64))


def test_batch_graph_survives_context_scratch_changes(ctx):
    """A recorded graph keeps its own sort scratch: a larger sort on the same
    context (which regrows the context's ping-pong buffers and look-back
    words), a trim (which frees them) and a merge in between leave every
    replay equal to the eager run."""
    import torch

    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    insts = W.batched_contractions(150, seed=15)
    ak, av, bk, bv = W.generate_operands(ctx, insts)
    b = Batch(ctx, insts)
    ek, ev, eoff = b.contract(ak, av, bk, bv)
    b.capture(ak, av, bk, bv)
    gk, gv, goff = b.replay()
    assert torch.equal(gk, ek) and torch.equal(gv.view(torch.int64), ev.view(torch.int64))
    n = max(4 * int(ak.shape[0]), 1 << 22)
    for step in ("bigger sort", "trim", "sort after trim"):
        if step == "trim":
            ctx.trim()
        else:
            k = torch.empty(n, dtype=torch.int64, device="cuda")
            v = torch.empty(n, dtype=torch.float64, device="cuda")
            ctx.testgen_distinct_keys(k, 1 << 40, 77)
            ctx.testgen_uniform(v, 78)
            ctx.sort_pairs(k, v, 1 << 40)
            torch.cuda.synchronize()
            assert bool((k[1:] > k[:-1]).all())
            n *= 2
        gk, gv, goff = b.replay()
        assert np.array_equal(goff, eoff), step
        assert torch.equal(gk, ek) and torch.equal(gv.view(torch.int64), ev.view(torch.int64)), step
    b.close()


def test_batch_graph_capacity_overflow(ctx):
    import libnt_amd
    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch

    insts = W.batched_contractions(50, seed=6)
    ak, av, bk, bv = W.generate_operands(ctx, insts)
    b = Batch(ctx, insts)
    n = b.capture(ak, av, bk, bv)
    assert n > 10
    b.capture(ak, av, bk, bv, capacity=n - 7)
    with pytest.raises(libnt_amd.LntError) as e:
        b.replay()
    assert e.value.status == 6


def test_batch_edge_cases(ctx, orc):
    import torch

    import libnt_amd
    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch, Instance

    insts = [
        Instance([4, 5, 6], "ikj", [6, 7], "jl", 0, 10),      # empty A
        Instance([4, 5, 6], "ikj", [6, 7], "jl", 12, 0),      # empty B
        Instance([1, 3], "ij", [3, 1], "jk", 2, 2),           # h = 1, bw = 1
        Instance([5, 4, 3], "iJk", [3, 4, 6], "kJm", 30, 40),  # inter index: 4 tabs
        Instance([9, 8, 7], "abc", [7, 8, 9], "cbd", 200, 150),  # contracted modes in opposite orders
        Instance([6, 6, 6, 6], "abcd", [6, 6, 6, 6], "cdef", 500, 600),
    ]
    b, ops, ck, cv, off = _run(ctx, insts)
    ck, cv = host(ck), host(cv)
    for i, x in enumerate(insts):
        ak, av, bk, bv = _oracle_operands(orc, insts, i)
        ident_a, ident_b = list(range(len(x.dims_a))), list(range(len(x.dims_b)))
        ek, ev, _ = orc.contract(ak, av, x.dims_a, ident_a, False, x.labels_a,
                                 bk, bv, x.dims_b, ident_b, False, x.labels_b)
        assert np.array_equal(ck[off[i]:off[i + 1]], ek), i
        assert orc.allclose_fuzzy(cv[off[i]:off[i + 1]], ev, TOL), i
    # empty batch
    e = torch.empty(0, dtype=torch.int64, device="cuda")
    ev = torch.empty(0, dtype=torch.float64, device="cuda")
    k, v, o = Batch(ctx, []).contract(e, ev, e, ev)
    assert k.shape[0] == 0 and list(o) == [0]
    # a key outside its instance's dimensionality
    ak, av, bk, bv = ops
    bad = ak.clone()
    bad[3] = insts[0].dimensionality_a * 1000
    with pytest.raises(libnt_amd.LntError) as err:
        b.contract(bad, av, bk, bv)
    assert err.value.status == 5


def test_batch_tall_lattices(ctx, orc):
    """An instance whose A lattice has more than 2^32 rows: the large-segment
    window sort keeps 64-bit keys for the whole batch (32-bit otherwise);
    results against the oracle either way."""
    from libnt_amd.batch import Instance

    tall = [Instance([1 << 17, 1 << 17, 4], "ijk", [4, 3], "kl", 20000, 12),  # h = 2^34: long columns
            Instance([6, 6, 6, 6], "abcd", [6, 6, 6, 6], "cdef", 500, 600)]
    short = [Instance([1 << 12, 1 << 12, 4], "ijk", [4, 3], "kl", 20000, 12), tall[1]]
    for insts in (tall, short):
        b, ops, ck, cv, off = _run(ctx, insts)
        ck, cv = host(ck), host(cv)
        for i, x in enumerate(insts):
            ak, av, bk, bv = _oracle_operands(orc, insts, i)
            ident_a, ident_b = list(range(len(x.dims_a))), list(range(len(x.dims_b)))
            ek, ev, _ = orc.contract(ak, av, x.dims_a, ident_a, False, x.labels_a,
                                     bk, bv, x.dims_b, ident_b, False, x.labels_b)
            assert np.array_equal(ck[off[i]:off[i + 1]], ek), i
            assert orc.allclose_fuzzy(cv[off[i]:off[i + 1]], ev, TOL), i


def test_batch_invalid_descriptions(ctx):
    import libnt_amd
    from libnt_amd.batch import Batch, Instance

    x = Instance([4, 5], "ij", [5, 6], "jk", 3, 3)
    x.lat_a = (4, 5, 2)  # inconsistent lattice shape
    with pytest.raises(libnt_amd.LntError) as e:
        Batch(ctx, [x])
    assert e.value.status == 7
    y = Instance([4, 5], "ij", [5, 6], "jk", 30, 3)  # nnz above the dimensionality
    with pytest.raises(libnt_amd.LntError) as e:
        Batch(ctx, [y])
    assert e.value.status == 1


def _linearity_checksums(insts, ak, av, bk, bv, a_off, b_off):
    """Per instance: sum_k colsum_A(k) * rowsum_B(k) and the bound
    sum_k colabs_A(k) * rowabs_B(k), computed from the operands alone."""
    import torch

    want = np.zeros(len(insts))
    bound = np.zeros(len(insts))
    for i, x in enumerate(insts):
        w = x.lat_a[1] * x.lat_a[2]

        def index_of(keys, dims, labels, contracted):
            # (contracted index, tab) of every key in default order, as one integer
            idx = torch.zeros_like(keys)
            stride, place = 1, {}
            for m, d in enumerate(dims):
                place[labels[m]] = (stride, d)
                stride *= d
            mult = 1
            for c in contracted:
                s, d = place[c]
                idx += ((keys // s) % d) * mult
                mult *= d
            return idx

        inner = [c for c in x.labels_a if c in x.labels_b and c.islower()]
        inter = [c for c in x.labels_a if c in x.labels_b and c.isupper()]
        ka = ak[a_off[i]:a_off[i + 1]]
        kb = bk[b_off[i]:b_off[i + 1]]
        ia = index_of(ka, x.dims_a, x.labels_a, inner + inter)
        ib = index_of(kb, x.dims_b, x.labels_b, inner + inter)
        ca = torch.zeros(w, dtype=torch.float64, device="cuda").index_add_(0, ia, av[a_off[i]:a_off[i + 1]])
        cb = torch.zeros(w, dtype=torch.float64, device="cuda").index_add_(0, ib, bv[b_off[i]:b_off[i + 1]])
        aa = torch.zeros(w, dtype=torch.float64, device="cuda").index_add_(0, ia, av[a_off[i]:a_off[i + 1]].abs())
        ab = torch.zeros(w, dtype=torch.float64, device="cuda").index_add_(0, ib, bv[b_off[i]:b_off[i + 1]].abs())
        want[i] = float((ca * cb).sum())
        bound[i] = float((aa * ab).sum())
    return want, bound


def test_c4_full_size_batch(ctx, orc):
    """The full batched mix: 10,000 instances, order 3-4, 1-2 contracted
    modes, dims log-U[8,128], nnz log-U[100,50k].  Checked by the recorded
    graph against the eager run, per-instance sortedness and bounds, the
    linearity checksum of every instance, and every 100th instance entry by
    entry against the oracle's full contraction."""
    import torch

    from libnt_amd import workloads as W
    from libnt_amd.batch import Batch, packed_offsets

    insts = W.batched_contractions(10000, seed=4)
    ak, av, bk, bv = W.generate_operands(ctx, insts)
    a_off = packed_offsets([x.nnz_a for x in insts])
    b_off = packed_offsets([x.nnz_b for x in insts])
    b = Batch(ctx, insts)
    nprod = b.nprod(ak, bk)
    print("c4: %d instances, %d + %d entries, %.3g products" % (len(insts), a_off[-1], b_off[-1], nprod.sum()),
          flush=True)
    ck, cv, off = b.contract(ak, av, bk, bv)
    nnz = np.diff(off)
    print("c4: nnz(C) = %d" % off[-1], flush=True)
    assert (nnz <= nprod).all()
    assert ((nnz > 0) == (nprod > 0)).all()
    # keys strictly ascending inside every instance (a step down is allowed
    # only where a new instance starts), first key >= 0, last key below C's
    # dimensionality
    starts = off[1:-1][(off[1:-1] > 0) & (off[1:-1] < off[-1])]
    down = []
    chunk = 1 << 28
    for lo in range(0, max(ck.shape[0] - 1, 0), chunk):
        seg = ck[lo:lo + chunk + 1]
        down.append(torch.nonzero(seg[1:] <= seg[:-1]).flatten().cpu().numpy() + lo + 1)
    down = np.concatenate(down) if down else np.zeros(0, dtype=np.int64)
    assert set(down.tolist()) <= set(starts.tolist())
    nz = np.nonzero(nnz)[0]
    first = ck[torch.from_numpy(off[nz]).cuda()].cpu().numpy()
    last = ck[torch.from_numpy(off[nz + 1] - 1).cuda()].cpu().numpy()
    dimc = np.array([math.prod(insts[i].dims_c) for i in nz], dtype=object)
    assert (first >= 0).all() and all(int(l) < int(d) for l, d in zip(last, dimc))
    # every 100th instance of the real mix against the oracle, entry by entry
    for i in range(0, len(insts), 100):
        x = insts[i]
        hak, hav, hbk, hbv = _oracle_operands(orc, insts, i)
        ek, ev, _ = orc.contract(hak, hav, x.dims_a, list(range(len(x.dims_a))), False, x.labels_a,
                                 hbk, hbv, x.dims_b, list(range(len(x.dims_b))), False, x.labels_b)
        got_k = ck[int(off[i]):int(off[i + 1])].cpu().numpy()
        got_v = cv[int(off[i]):int(off[i + 1])].cpu().numpy()
        assert np.array_equal(got_k, ek), "instance %d keys" % i
        assert orc.allclose_fuzzy(got_v, ev, TOL), "instance %d values" % i
    # per-instance sums against the linearity checksum
    fp_eager = _fingerprints(ck, cv, off)
    got = fp_eager[2]
    want, bound = _linearity_checksums(insts, ak, av, bk, bv, a_off, b_off)
    assert np.all(np.abs(got - want) <= 1e-10 * bound + 1e-300)
    # the recorded graph reproduces the eager result bit for bit (compared by
    # per-instance fingerprints: the eager result is freed first)
    del ck, cv
    torch.cuda.empty_cache()
    n = b.capture(ak, av, bk, bv)
    assert n == off[-1]
    gk, gv, goff = b.replay()
    assert np.array_equal(goff, off)
    fp_graph = _fingerprints(gk, gv, goff)
    assert np.array_equal(fp_eager[0], fp_graph[0])  # key words
    assert np.array_equal(fp_eager[1], fp_graph[1])  # value bit patterns


def _fingerprints(ck, cv, off):
    """Per instance: wrapping sums of the key words and of the value bit
    patterns, and the floating-point sum of the values."""
    import torch

    ks, vs, fs = [], [], []
    vb = cv.view(torch.int64)
    for i in range(len(off) - 1):
        a, b = int(off[i]), int(off[i + 1])
        ks.append(ck[a:b].sum())
        vs.append(vb[a:b].sum())
        fs.append(cv[a:b].sum())
    return (torch.stack(ks).cpu().numpy(), torch.stack(vs).cpu().numpy(), torch.stack(fs).cpu().numpy())
