"""GPU tests of single contractions by strategy and of the learned cost model
(include/lnt_dispatch.h), through the C-ABI:
  * every strategy returns the oracle's contraction (keys bit-exact, values
    within 1e-12 under isEqualFuzzy, FunctionUtil.h:688-698) for operands in
    every sort state, and the sort-accumulate strategy equals the front end's
    product bit for bit;
  * the model's forward pass and its loss gradient (FP32 kernels) against the
    FP64 restatement in the oracle;
  * training by the GPU kernels lowers the loss;
  * with the committed timing set: the trained model's choices against the
    restated polyalgorithm on held-out instances (regret)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libnt_amd", "data")


def host(t):
    return t.cpu().numpy()


def _specs():
    from libnt_amd import dispatch

    specs = dispatch.mixed_specs(12, seed=5, max_nnz=4000)
    # extra shapes: inter index, pure outer, pure Hadamard, two contracted modes in opposite orders
    specs.append(dispatch.Spec("x", [6, 5, 7], "iJk", [7, 5, 9], "kJm", 120, 150, "ab", "unsorted"))
    specs.append(dispatch.Spec("x", [30, 40], "ij", [50, 20], "kl", 200, 300, "default", "ba"))
    specs.append(dispatch.Spec("x", [9, 8, 7], "IJK", [7, 9, 8], "KIJ", 200, 250, "unsorted", "default"))
    specs.append(dispatch.Spec("x", [9, 8, 7], "abc", [8, 9, 6], "bad", 300, 200, "ba", "ab"))
    return specs


def _oracle(orc, A, la, B, lb):
    return orc.contract(host(A.keys), host(A.vals), list(A.dims), A.seq, A.is_sorted, la,
                        host(B.keys), host(B.vals), list(B.dims), B.seq, B.is_sorted, lb)


# Accumulators whose sums are the same chain of additions (each row in
# ascending B row order): one orientation gives the same bits with any of them.
_ORDERED = (0, 2, 3, 4)  # sort, dense, esc, mtt


def test_every_strategy_matches_oracle(ctx, orc):
    import torch

    from libnt_amd import dispatch

    for i, sp in enumerate(_specs()):
        A, B = dispatch.materialize(ctx, sp, 100 + i)
        ek, ev, edims = _oracle(orc, A, sp.la, B, sp.lb)
        # csc_times's sparse accumulator sums in ascending B row order: the
        # bits of the A*B-oriented sort, dense, esc and mtt strategies.
        sk, sv, _ = orc.contract(host(A.keys), host(A.vals), list(A.dims), A.seq, A.is_sorted, sp.la,
                                 host(B.keys), host(B.vals), list(B.dims), B.seq, B.is_sorted, sp.lb, method="spa")
        assert np.array_equal(sk, ek)
        ref = None
        for s in range(dispatch.NUM_STRATEGIES):
            print("spec", i, sp.kind, sp.la, sp.lb, "strategy", dispatch.strategy_name(s), flush=True)
            C, labels, used = dispatch.contract(ctx, A, sp.la, B, sp.lb, s)
            assert used == s
            assert list(C.dims) == list(edims)
            assert np.array_equal(host(C.keys), ek), (i, s)
            assert orc.allclose_fuzzy(host(C.vals), ev, TOL), (i, s)
            if dispatch.strategy_acc(s) in _ORDERED and not s & dispatch.STRAT_SWAP:
                assert np.array_equal(host(C.vals).view(np.int64), sv.view(np.int64)), (i, s)
            if s == 0:
                ref = C
        # the front end's product (sort-accumulate, A*B orientation) bit for bit
        from libnt_amd.sparse import SparseMIA

        A2 = SparseMIA(A.dims, A.keys.clone(), A.vals.clone(), A.seq, A.is_sorted, ctx)
        B2 = SparseMIA(B.dims, B.keys.clone(), B.vals.clone(), B.seq, B.is_sorted, ctx)
        F = (A2(sp.la) * B2(sp.lb)).mia
        assert torch.equal(F.keys, ref.keys)
        assert torch.equal(F.vals.view(torch.int64), ref.vals.view(torch.int64))


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_random_contractions_every_strategy(ctx, orc, seed):
    """Wider sweep: 15 instances per seed from the three families (up to 30k
    entries per operand, operands in every sort state), every strategy
    against the oracle."""
    from libnt_amd import dispatch

    for i, sp in enumerate(dispatch.mixed_specs(15, seed=seed, max_nnz=30000)):
        A, B = dispatch.materialize(ctx, sp, 1000 * seed + i)
        ek, ev, edims = _oracle(orc, A, sp.la, B, sp.lb)
        for s in range(dispatch.NUM_STRATEGIES):
            print("seed", seed, "spec", i, sp.kind, sp.la, sp.lb, sp.nnz_a, sp.nnz_b, "strategy", s, flush=True)
            C, labels, used = dispatch.contract(ctx, A, sp.la, B, sp.lb, s)
            assert list(C.dims) == list(edims)
            assert np.array_equal(host(C.keys), ek), (seed, i, s)
            assert orc.allclose_fuzzy(host(C.vals), ev, TOL), (seed, i, s)


def test_auto_without_model_is_the_heuristic(ctx):
    from libnt_amd import dispatch

    for i, sp in enumerate(_specs()[:8]):
        A, B = dispatch.materialize(ctx, sp, 200 + i)
        _, _, used = dispatch.contract(ctx, A, sp.la, B, sp.lb, dispatch.STRAT_AUTO)
        assert used == dispatch.heuristic(A, sp.la, B, sp.lb)


def test_time_strategies(ctx):
    from libnt_amd import dispatch

    sp = _specs()[0]
    A, B = dispatch.materialize(ctx, sp, 300)
    ms = dispatch.time_strategies(ctx, A, sp.la, B, sp.lb, dispatch.ALL_STRATEGIES, 1)
    assert ms.shape == (dispatch.NUM_STRATEGIES,) and (ms > 0).all()
    ms = dispatch.time_strategies(ctx, A, sp.la, B, sp.lb, 0b101 | (1 << 19), 1)
    assert ms[0] > 0 and ms[2] > 0 and ms[19] > 0 and ms[1] == -1 and ms[3] == -1 and ms[18] == -1


def _random_model(ctx, orc, seed):
    from libnt_amd import dispatch

    rng = np.random.default_rng(seed)
    m = dispatch.Model(ctx, seed)
    p = m.params.astype(np.float64)
    nw = orc.mlp_weight_count()
    p[:nw] += rng.normal(0, 0.05, nw)          # biases non-zero too
    p[nw:nw + 18] = rng.normal(0, 1, 18)
    p[nw + 18:] = rng.uniform(0.5, 2.0, 18)
    m.params = p.astype(np.float32)
    return m, m.params.astype(np.float64)


def test_model_forward_matches_oracle(ctx, orc):
    m, P = _random_model(ctx, orc, 11)
    X = np.random.default_rng(1).normal(0, 2, (300, 18)).astype(np.float32)
    got = m.predict(X)
    want = orc.mlp_forward(P, X.astype(np.float64))
    assert np.allclose(got, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())


@pytest.mark.parametrize("n", [1, 37, 700])
def test_model_gradient_matches_oracle(ctx, orc, n):
    from libnt_amd import dispatch

    m, P = _random_model(ctx, orc, 12 + n)
    rng = np.random.default_rng(n)
    X = rng.normal(0, 2, (n, 18)).astype(np.float32)
    T = rng.normal(0, 2, (n, dispatch.NUM_STRATEGIES)).astype(np.float32)
    T[rng.random((n, dispatch.NUM_STRATEGIES)) < 0.25] = np.nan
    if not np.isfinite(T).any():
        T[0, 0] = 1.0
    loss, g = m.gradient(X, T)
    wl, wg = orc.mlp_loss_grad(P, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - wl) <= 1e-4 * abs(wl)
    scale = np.abs(wg).max()
    assert np.allclose(g, wg, rtol=1e-3, atol=1e-4 * scale)


def test_model_training_lowers_loss(ctx):
    from libnt_amd import dispatch

    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (2000, 18)).astype(np.float32)
    W = rng.normal(0, 0.5, (18, dispatch.NUM_STRATEGIES))
    T = (np.tanh(X @ W) + 0.1 * X @ np.eye(18, dispatch.NUM_STRATEGIES)).astype(np.float32)
    m = dispatch.Model(ctx, 3)
    l0, _ = m.gradient(X, T)
    l1 = m.train(X, T, epochs=60, batch=128, lr=3e-3, seed=1)
    assert l1 < 0.1 * l0, (l0, l1)
    # training is deterministic
    m2 = dispatch.Model(ctx, 3)
    l2 = m2.train(X, T, epochs=60, batch=128, lr=3e-3, seed=1)
    assert l2 == l1 and np.array_equal(m.params, m2.params)


def _timing_set():
    path = os.path.join(DATA, "dispatch_timings.npz")
    if not os.path.exists(path):
        pytest.skip("no committed timing set")
    return np.load(path, allow_pickle=False)


def test_learned_dispatch_beats_the_heuristic_on_held_out_instances(ctx):
    """Trained on the committed training split by the GPU kernels, the model's
    choices on the held-out split cost less in total, and have lower mean
    regret, than the restated polyalgorithm's (timings measured on MI355X)."""
    from libnt_amd import dispatch

    d = _timing_set()
    Xtr, Ttr, Xte, Tte, Hte = d["X_train"], d["T_train"], d["X_test"], d["T_test"], d["H_test"]
    m = dispatch.Model(ctx, 1)
    m.train(Xtr, Ttr, epochs=int(d["epochs"]), batch=256, lr=float(d["lr"]), seed=7)
    ms = np.exp2(Tte.astype(np.float64))
    learned = dispatch.choose(m.predict(Xte))
    t_learned = ms[np.arange(len(ms)), learned].sum()
    t_heur = ms[np.arange(len(ms)), Hte].sum()
    r_learned = dispatch.regret(ms, learned).mean()
    r_heur = dispatch.regret(ms, Hte).mean()
    print("held-out: learned %.1f ms (regret %.3f), heuristic %.1f ms (regret %.3f), best %.1f ms"
          % (t_learned, r_learned, t_heur, r_heur, ms.min(axis=1).sum()), flush=True)
    assert t_learned < t_heur
    assert r_learned < r_heur


def test_shipped_model_drives_auto(ctx):
    """The committed weights load, and AUTO runs the strategy the model
    predicts fastest, with the same result as the fixed strategies."""
    from libnt_amd import dispatch

    path = os.path.join(DATA, "dispatch_model.npy")
    if not os.path.exists(path):
        pytest.skip("no committed model")
    m = dispatch.Model(ctx, 1)
    m.load(path)
    m.attach()
    try:
        for i, sp in enumerate(_specs()[:6]):
            A, B = dispatch.materialize(ctx, sp, 400 + i)
            f = dispatch.features(A, sp.la, B, sp.lb)
            want = int(np.argmin(m.predict(f[None, :])[0]))
            C, _, used = dispatch.contract(ctx, A, sp.la, B, sp.lb, dispatch.STRAT_AUTO)
            assert used == want
            C0, _, _ = dispatch.contract(ctx, A, sp.la, B, sp.lb, 0)
            assert np.array_equal(host(C.keys), host(C0.keys))
    finally:
        from libnt_amd import check, lib

        check(lib().lnt_context_set_model(ctx.handle, None))
