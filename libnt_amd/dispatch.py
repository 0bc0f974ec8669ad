"""Single contractions by strategy and the learned dispatch
(``include/lnt_dispatch.h``).

``contract`` runs ``C = A(la) * B(lb)`` -- the reference's expression
(MIA_Expr.h:339-419 unfold, SparseLatticeBase.h:1574 product,
SparseMIA.h:362-372 refold) -- by one of ``NUM_STRATEGIES`` strategies
(orientation x accumulator x radix width) with identical results.
``heuristic`` is the reference's fixed rule (sparse_times_polyalgorithm.m);
``Model`` is the learned cost model, trained on the GPU by the library's own
kernels on timings that ``collect_timings`` measures.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import MEM_DEVICE, _i64p, _vp, check, lib

MAX_ORDER = 16
NUM_FEATURES = 18
STRAT_SWAP, STRAT_RADIX6 = 1, 2
ACC_SORT, ACC_HASH, ACC_DENSE, ACC_ESC, ACC_MTT = range(5)
ACC_NAMES = ("sort", "hash", "dense", "esc", "mtt")
NUM_STRATEGIES = 4 * len(ACC_NAMES)
ALL_STRATEGIES = (1 << NUM_STRATEGIES) - 1
STRAT_AUTO = -1
HIDDEN = 64
FEATURE_NAMES = [
    "log2_nnz_a", "log2_nnz_b", "log2_h", "log2_w", "log2_bw", "log2_d", "order_a", "order_b", "n_contracted",
    "n_inter", "log2_est_products", "log2_est_nnz_c", "a_ready_ab", "b_ready_ab", "a_ready_ba", "b_ready_ba",
    "log2_dim_a", "log2_dim_b",
]


def strategy(acc: int, radix6: bool = False, swap: bool = False) -> int:
    """``LNT_STRATEGY(acc, radix6, swap)`` of include/lnt_dispatch.h."""
    return (acc << 2) | (STRAT_RADIX6 if radix6 else 0) | (STRAT_SWAP if swap else 0)


def strategy_acc(s: int) -> int:
    return s >> 2


def strategy_name(s: int) -> str:
    return "%s/%s/r%d" % ("BA" if s & STRAT_SWAP else "AB", ACC_NAMES[s >> 2], 6 if s & STRAT_RADIX6 else 8)


class Operand(ctypes.Structure):
    """``lnt_operand``."""

    _fields_ = [
        ("order", ctypes.c_int32),
        ("dims", ctypes.c_int64 * MAX_ORDER),
        ("seq", ctypes.c_int32 * MAX_ORDER),
        ("labels", ctypes.c_char * (MAX_ORDER + 1)),
        ("sorted", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("keys", ctypes.c_void_p),
        ("vals", ctypes.c_void_p),
    ]


class ResultShape(ctypes.Structure):
    """``lnt_result_shape``."""

    _fields_ = [
        ("order", ctypes.c_int32),
        ("dims", ctypes.c_int64 * MAX_ORDER),
        ("labels", ctypes.c_char * (MAX_ORDER + 1)),
    ]


def operand(mia, labels: str) -> Operand:
    """Descriptor of a ``libnt_amd.sparse.SparseMIA`` labelled by ``labels``."""
    if len(labels) != mia.order:
        raise ValueError("one label per mode is required")
    o = Operand()
    o.order = mia.order
    for m in range(mia.order):
        o.dims[m] = mia.dims[m]
        o.seq[m] = mia.seq[m]
    o.labels = labels.encode()
    o.sorted = int(bool(mia.is_sorted))
    o.nnz = mia.nnz
    o.keys = mia.keys.data_ptr() if mia.nnz else None
    o.vals = mia.vals.data_ptr() if mia.nnz else None
    return o


def contract(ctx, a, la: str, b, lb: str, strategy: int = STRAT_AUTO):
    """C = A(la) * B(lb); returns (SparseMIA C, labels of C, strategy used)."""
    import torch

    from .sparse import SparseMIA

    oa, ob = operand(a, la), operand(b, lb)
    ck, cv = _vp(), _vp()
    n = ctypes.c_int64()
    shape = ResultShape()
    used = ctypes.c_int()
    check(lib().lnt_contract(ctx.handle, ctypes.byref(oa), ctypes.byref(ob), int(strategy), ctypes.byref(ck),
                             ctypes.byref(cv), ctypes.byref(n), ctypes.byref(shape), ctypes.byref(used)))
    try:
        keys = torch.empty(max(n.value, 1), dtype=torch.int64, device="cuda")
        vals = torch.empty(max(n.value, 1), dtype=torch.float64, device="cuda")
        check(lib().lnt_copy(ctx.handle, keys.data_ptr(), ck, 8 * n.value))
        check(lib().lnt_copy(ctx.handle, vals.data_ptr(), cv, 8 * n.value))
    finally:
        lib().lnt_free(ctx.handle, ck, MEM_DEVICE)
        lib().lnt_free(ctx.handle, cv, MEM_DEVICE)
    dims = [shape.dims[m] for m in range(shape.order)]
    c = SparseMIA(dims, keys[: n.value], vals[: n.value], None, True, ctx)
    return c, shape.labels.decode(), used.value


def features(a, la: str, b, lb: str) -> np.ndarray:
    oa, ob = operand(a, la), operand(b, lb)
    f = np.zeros(NUM_FEATURES, dtype=np.float32)
    check(lib().lnt_contract_features(ctypes.byref(oa), ctypes.byref(ob),
                                      f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return f


def heuristic(a, la: str, b, lb: str) -> int:
    oa, ob = operand(a, la), operand(b, lb)
    s = ctypes.c_int()
    check(lib().lnt_contract_heuristic(ctypes.byref(oa), ctypes.byref(ob), ctypes.byref(s)))
    return s.value


BATCH_POLICIES = ("sort", "dense", "heuristic", "learned")


def batch_accumulators(pairs, policy: str, model: Optional["Model"] = None) -> np.ndarray:
    """Per-instance accumulators of a batch (``Batch.set_accumulators``) from
    the instances' operand descriptors ``pairs`` = [(Operand A, Operand B)]:
    "sort" the batch's own merge/sort paths everywhere, "dense" the dense
    accumulator wherever it fits, "heuristic" the accumulator class of the
    restated polyalgorithm (``lnt_contract_heuristic``), "learned" the
    accumulator of the model's fastest predicted strategy in the batch's
    orientation (A*B)."""
    n = len(pairs)
    if policy == "sort":
        return np.full(n, ACC_SORT, dtype=np.int32)
    if policy == "dense":
        return np.full(n, ACC_DENSE, dtype=np.int32)
    if policy == "heuristic":
        out = np.zeros(n, dtype=np.int32)
        s = ctypes.c_int()
        for i, (oa, ob) in enumerate(pairs):
            check(lib().lnt_contract_heuristic(ctypes.byref(oa), ctypes.byref(ob), ctypes.byref(s)))
            out[i] = strategy_acc(s.value)
        return out
    if policy == "learned":
        if model is None:
            raise ValueError("the learned policy needs a model")
        X = np.zeros((n, NUM_FEATURES), dtype=np.float32)
        for i, (oa, ob) in enumerate(pairs):
            check(lib().lnt_contract_features(ctypes.byref(oa), ctypes.byref(ob), _fp(X[i])))
        pred = model.predict(X) if n else np.zeros((0, NUM_STRATEGIES), dtype=np.float32)
        ab = [s for s in range(NUM_STRATEGIES) if not s & STRAT_SWAP]
        return np.array([strategy_acc(ab[int(j)]) for j in np.argmin(pred[:, ab], axis=1)], dtype=np.int32)
    raise ValueError("unknown policy %r (one of %s)" % (policy, ", ".join(BATCH_POLICIES)))


def time_strategies(ctx, a, la: str, b, lb: str, mask: int = ALL_STRATEGIES, reps: int = 1) -> np.ndarray:
    """Wall ms of each strategy in mask (min over reps after a warm-up); -1
    where not measured."""
    oa, ob = operand(a, la), operand(b, lb)
    ms = np.zeros(NUM_STRATEGIES)
    check(lib().lnt_contract_time(ctx.handle, ctypes.byref(oa), ctypes.byref(ob), int(mask), int(reps),
                                  ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return ms


def param_count() -> int:
    return int(lib().lnt_model_param_count())


class Model:
    """The learned cost model: predicted log2(ms) of every strategy."""

    def __init__(self, ctx, seed: int = 1):
        self.ctx = ctx
        self._h = _vp()
        check(lib().lnt_model_create(ctx.handle, int(seed), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().lnt_model_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(param_count(), dtype=np.float32)
        check(lib().lnt_model_get_params(self._h, _fp(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (param_count(),):
            raise ValueError("expected %d parameters" % param_count())
        check(lib().lnt_model_set_params(self._h, _fp(p)))

    def train(self, X, T, epochs=200, batch=256, lr=1e-3, seed=7) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = ctypes.c_float()
        check(lib().lnt_model_train(self._h, _fp(X), _fp(T), X.shape[0], int(epochs), int(batch), float(lr),
                                    int(seed), ctypes.byref(loss)))
        return loss.value

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], NUM_STRATEGIES), dtype=np.float32)
        check(lib().lnt_model_predict(self._h, _fp(X), X.shape[0], _fp(out)))
        return out

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(param_count() - 2 * NUM_FEATURES, dtype=np.float32)
        loss = ctypes.c_float()
        check(lib().lnt_model_gradient(self._h, _fp(X), _fp(T), X.shape[0], _fp(g), ctypes.byref(loss)))
        return loss.value, g

    def attach(self):
        """Use this model for STRAT_AUTO in its context."""
        check(lib().lnt_context_set_model(self.ctx.handle, self._h))

    def save(self, path):
        np.save(path, self.params)

    def load(self, path):
        self.params = np.load(path, allow_pickle=False)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# ------------------------------------------------------------ training data
def choose(pred_log_ms: np.ndarray) -> np.ndarray:
    return np.argmin(pred_log_ms, axis=1)


def regret(times_ms: np.ndarray, chosen: np.ndarray) -> np.ndarray:
    """t_chosen / t_best - 1 per instance (times_ms: n x strategies)."""
    best = times_ms.min(axis=1)
    return times_ms[np.arange(len(chosen)), chosen] / best - 1.0


class Spec:
    """Description of one instance of the mixed dispatch workload."""

    def __init__(self, kind, dims_a, la, dims_b, lb, nnz_a, nnz_b, state_a, state_b, zipf_a=(), zipf_b=()):
        self.kind = kind
        self.dims_a, self.la, self.dims_b, self.lb = list(dims_a), la, list(dims_b), lb
        self.nnz_a, self.nnz_b = int(nnz_a), int(nnz_b)
        self.state_a, self.state_b = state_a, state_b  # "unsorted", "default", "ab", "ba"
        self.zipf_a, self.zipf_b = tuple(zipf_a), tuple(zipf_b)  # modes with Zipf(1.1) coordinates


def _logu(rng, lo, hi):
    return int(min(hi, math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi + 1))))))


def mixed_specs(count: int, seed: int, max_nnz: int = 200000) -> List[Spec]:
    """Instances drawn from three families: the batched mix (order 3-4, 1-2
    contracted modes, dims log-U[8,128]); order-4 tensors with dims up to 256
    (the sort/merge family) contracting 1-2 modes; the skewed family A(i,j,k) *
    B(k,l,m) with Zipf(1.1) i and l.  Each operand arrives unsorted, sorted in
    the default order, or sorted in the lattice order of either orientation."""
    from .workloads import CONTRACTED, OUTER_A, OUTER_B

    specs = []
    states = ["unsorted", "default", "ab", "ba"]
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        kind = ("batched", "order4", "skewed")[i % 3]
        if kind == "skewed":
            di, dj, dl = (1 << int(rng.integers(8, 13)) for _ in range(3))
            dk, dm = 1 << int(rng.integers(5, 10)), 1 << int(rng.integers(4, 7))
            dims_a, la, dims_b, lb = [di, dj, dk], "ijk", [dk, dl, dm], "klm"
            zipf_a, zipf_b = (0,), (1,)
        else:
            if kind == "batched":
                oa, ob = (int(x) for x in rng.integers(3, 5, size=2))
                lo, hi = 8, 128
            else:
                oa = ob = 4
                lo, hi = 16, 256
            c = int(rng.integers(1, 3))
            dc = [_logu(rng, lo, hi) for _ in range(c)]
            la_l = list(CONTRACTED[:c]) + list(OUTER_A[: oa - c])
            lb_l = list(CONTRACTED[:c]) + list(OUTER_B[: ob - c])
            da = dc + [_logu(rng, lo, hi) for _ in range(oa - c)]
            db = dc + [_logu(rng, lo, hi) for _ in range(ob - c)]
            pa, pb = rng.permutation(oa), rng.permutation(ob)
            la = "".join(la_l[p] for p in pa)
            lb = "".join(lb_l[p] for p in pb)
            dims_a = [da[p] for p in pa]
            dims_b = [db[p] for p in pb]
            zipf_a = zipf_b = ()
        na = min(_logu(rng, 1000, max_nnz), math.prod(dims_a) // 2)
        nb = min(_logu(rng, 1000, max_nnz), math.prod(dims_b) // 2)
        sa, sb = states[int(rng.integers(0, 4))], states[int(rng.integers(0, 4))]
        specs.append(Spec(kind, dims_a, la, dims_b, lb, na, nb, sa, sb, zipf_a, zipf_b))
    return specs


def batch_specs(count: int, seed: int) -> List[Spec]:
    """Instances of the batched contraction mix (workloads.batched_contractions:
    nnz log-U[100, 50k], order 3-4, 1-2 contracted modes, dims log-U[8,128]),
    both operands unsorted, as the per-instance calls over that mix see them."""
    from .workloads import batched_contractions

    return [Spec("batch", x.dims_a, x.labels_a, x.dims_b, x.labels_b, x.nnz_a, x.nnz_b, "unsorted", "unsorted")
            for x in batched_contractions(count, seed=seed)]


def _zipf_index(rng, n, size, s=1.1):
    """Zipf(s) over [0, n) by inverse CDF."""
    w = 1.0 / np.arange(1, n + 1) ** s
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, rng.random(size)), n - 1)


def materialize(ctx, spec: Spec, seed: int):
    """(A, B) SparseMIA operands of a spec on the GPU, in the spec's sort state."""
    import torch

    from .batch import lattice_plan
    from .sparse import SparseMIA

    def make(dims, n, zipf, kseed, vseed):
        if zipf:
            rng = np.random.default_rng([seed, kseed])
            subs = []
            for m, d in enumerate(dims):
                subs.append(_zipf_index(rng, d, 3 * n) if m in zipf else rng.integers(0, d, 3 * n))
            key = np.zeros(3 * n, dtype=np.int64)
            stride = 1
            for m, d in enumerate(dims):
                key += subs[m].astype(np.int64) * stride
                stride *= d
            key = np.unique(key)
            key = rng.permutation(key)[:n]
            keys = torch.from_numpy(key).cuda()
        else:
            keys = torch.empty(max(n, 1), dtype=torch.int64, device="cuda")
            ctx.testgen_distinct_keys(keys, math.prod(dims), kseed)
            keys = keys[:n]
        vals = torch.empty(max(keys.shape[0], 1), dtype=torch.float64, device="cuda")
        ctx.testgen_uniform(vals, vseed)
        return SparseMIA(dims, keys.contiguous(), vals[: keys.shape[0]].contiguous(), ctx=ctx)

    A = make(spec.dims_a, spec.nnz_a, spec.zipf_a, seed * 4 + 1, seed * 4 + 2)
    B = make(spec.dims_b, spec.nnz_b, spec.zipf_b, seed * 4 + 3, seed * 4 + 4)
    ab = lattice_plan(spec.dims_a, spec.la, spec.dims_b, spec.lb)
    ba = lattice_plan(spec.dims_b, spec.lb, spec.dims_a, spec.la)
    for X, state, seq_ab, seq_ba in ((A, spec.state_a, ab[0], ba[1]), (B, spec.state_b, ab[1], ba[0])):
        if state == "default":
            X.sort()
        elif state == "ab":
            X.sort(seq_ab)
        elif state == "ba":
            X.sort(seq_ba)
    return A, B


def collect_timings(ctx, specs: Sequence[Spec], seed: int, reps: int = 1, log=None):
    """Features (n x 18, float32) and log2 ms targets (n x NUM_STRATEGIES,
    float32) of the specs, every strategy measured."""
    X = np.zeros((len(specs), NUM_FEATURES), dtype=np.float32)
    T = np.zeros((len(specs), NUM_STRATEGIES), dtype=np.float32)
    H = np.zeros(len(specs), dtype=np.int32)
    for i, sp in enumerate(specs):
        A, B = materialize(ctx, sp, seed + i)
        X[i] = features(A, sp.la, B, sp.lb)
        H[i] = heuristic(A, sp.la, B, sp.lb)
        ms = time_strategies(ctx, A, sp.la, B, sp.lb, ALL_STRATEGIES, reps)
        T[i] = np.log2(np.maximum(ms, 1e-3))
        if log and (i + 1) % 200 == 0:
            log("%d/%d instances timed" % (i + 1, len(specs)))
    return X, T, H
