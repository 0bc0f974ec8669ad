"""CPU tests of the dispatch layer (include/lnt_dispatch.h) and its oracle:
the FP64 model restatement against central finite differences, the ctypes
structs against the C compiler's layout, the features against their closed
form, and the restated polyalgorithm (sparse_times_polyalgorithm.m) on
branches chosen by hand.  No GPU call is made: features and the heuristic
read descriptors only."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _random_params(orc, rng, scale=0.3):
    p = rng.normal(0, scale, orc.mlp_param_count())
    nw = orc.mlp_weight_count()
    p[nw:nw + 18] = rng.normal(0, 1, 18)           # mean
    p[nw + 18:] = rng.uniform(0.5, 2.0, 18)        # inv_std
    return p


def test_oracle_model_gradient_matches_finite_differences(orc):
    rng = np.random.default_rng(3)
    P = _random_params(orc, rng)
    X = rng.normal(0, 2, (7, 18))
    S = orc.MLP_OUTPUTS
    T = rng.normal(0, 1, (7, S))
    T[rng.random((7, S)) < 0.3] = np.nan  # unmeasured entries
    loss, g = orc.mlp_loss_grad(P, X, T)
    nw = orc.mlp_weight_count()
    idx = list(rng.choice(nw, 300, replace=False)) + [nw - 1, 0, 18 * 64 + 3]
    h = 1e-6
    for p in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        fd = (orc.mlp_loss_grad(Pp, X, T)[0] - orc.mlp_loss_grad(Pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(g[p])), (p, fd, g[p])


def test_oracle_model_loss_closed_form(orc):
    """All weights zero: y = b3, so the loss is the mean squared difference
    between b3 and the measured targets, and only b3 gets a gradient."""
    rng = np.random.default_rng(4)
    P = np.zeros(orc.mlp_param_count())
    nw = orc.mlp_weight_count()
    P[nw + 18:] = 1.0
    S = orc.MLP_OUTPUTS
    b3 = rng.normal(0, 1, S)
    P[nw - S:nw] = b3
    X = rng.normal(0, 1, (5, 18))
    T = rng.normal(0, 1, (5, S))
    T[0, 3] = np.nan
    loss, g = orc.mlp_loss_grad(P, X, T)
    m = np.isfinite(T)
    d = np.where(m, b3[None, :] - T, 0.0)
    assert math.isclose(loss, (d ** 2).sum() / m.sum(), rel_tol=1e-14)
    assert np.allclose(g[nw - S:], 2 * d.sum(axis=0) / m.sum(), rtol=1e-13, atol=0)
    assert not np.any(g[: nw - S])
    assert np.allclose(orc.mlp_forward(P, X), np.tile(b3, (5, 1)))


def test_model_parameter_count_matches_oracle(orc):
    from libnt_amd import dispatch

    assert dispatch.param_count() == orc.mlp_param_count()
    assert orc.MLP_OUTPUTS == dispatch.NUM_STRATEGIES


def test_struct_layouts_match_c(tmp_path):
    """ctypes mirrors of lnt_operand, lnt_result_shape and lnt_contraction
    have the C compiler's size and field offsets."""
    from libnt_amd import batch, dispatch

    src = tmp_path / "layout.c"
    src.write_text("""
#include <stddef.h>
#include <stdio.h>
#include "lnt_dispatch.h"
#include "lnt_batch.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(lnt_operand), offsetof(lnt_operand, dims),
         offsetof(lnt_operand, seq), offsetof(lnt_operand, labels), offsetof(lnt_operand, sorted),
         offsetof(lnt_operand, nnz), offsetof(lnt_operand, keys), offsetof(lnt_operand, vals));
  printf("%zu %zu %zu\\n", sizeof(lnt_result_shape), offsetof(lnt_result_shape, dims), offsetof(lnt_result_shape, labels));
  printf("%zu %zu %zu %zu\\n", sizeof(lnt_contraction), offsetof(lnt_contraction, cur_seq_a),
         offsetof(lnt_contraction, lat_a), offsetof(lnt_contraction, nnz_b));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    lines = subprocess.check_output([str(exe)]).decode().split("\n")
    O, R, C = dispatch.Operand, dispatch.ResultShape, batch.Contraction
    assert [int(x) for x in lines[0].split()] == [ctypes.sizeof(O), O.dims.offset, O.seq.offset, O.labels.offset,
                                                   O.sorted.offset, O.nnz.offset, O.keys.offset, O.vals.offset]
    assert [int(x) for x in lines[1].split()] == [ctypes.sizeof(R), R.dims.offset, R.labels.offset]
    assert [int(x) for x in lines[2].split()] == [ctypes.sizeof(C), C.cur_seq_a.offset, C.lat_a.offset,
                                                   C.nnz_b.offset]


def _op(dims, labels, nnz, seq=None, is_sorted=False):
    from libnt_amd import dispatch

    o = dispatch.Operand()
    o.order = len(dims)
    for m, d in enumerate(dims):
        o.dims[m] = d
        o.seq[m] = m if seq is None else seq[m]
    o.labels = labels.encode()
    o.sorted = int(is_sorted)
    o.nnz = nnz
    o.keys = 64 if nnz else None  # descriptors only: never dereferenced
    o.vals = 64 if nnz else None
    return o


def _features(a, b):
    from libnt_amd import check, lib

    f = np.zeros(18, dtype=np.float32)
    check(lib().lnt_contract_features(ctypes.byref(a), ctypes.byref(b), f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return f


def _heuristic(a, b):
    from libnt_amd import check, lib

    s = ctypes.c_int()
    check(lib().lnt_contract_heuristic(ctypes.byref(a), ctypes.byref(b), ctypes.byref(s)))
    return s.value


def test_features_closed_form():
    """A(i,K,j) [8,4,16] * B(j,K,l) [16,4,32]: h=8, w=16, d=4, bw=32."""
    a = _op([8, 4, 16], "iKj", 300, seq=[0, 2, 1], is_sorted=True)   # seq = [i, j, K]: A's AB lattice order
    b = _op([16, 4, 32], "jKl", 900)
    f = _features(a, b)
    nprod = 300 * 900 / (16 * 4)
    cells = 8 * 32 * 4
    nnzc = cells * -math.expm1(-nprod / cells)
    want = [math.log2(301), math.log2(901), 3, 4, 5, 2, 3, 3, 1, 1, math.log2(nprod + 1), math.log2(nnzc + 1),
            1, 0, 0, 0, math.log2(8 * 4 * 16), math.log2(16 * 4 * 32)]
    assert np.allclose(f, want, rtol=1e-6, atol=1e-6)
    # B sorted in the swapped orientation's first-operand order [l, j, K]
    b2 = _op([16, 4, 32], "jKl", 900, seq=[2, 0, 1], is_sorted=True)
    f2 = _features(a, b2)
    assert list(f2[12:16]) == [1, 0, 0, 1]


def test_heuristic_branches():
    """The orientation (`transposed`) and algorithm class of each branch of
    the polyalgorithm, mapped to accumulators: CSC/DCSC (1, 2) -> dense,
    operator*/CSCNA (0, 3) -> sort, outer product (4) -> esc."""
    from libnt_amd import dispatch

    SW = dispatch.STRAT_SWAP
    DENSE, SORT, ESC = (dispatch.strategy(a) for a in (dispatch.ACC_DENSE, dispatch.ACC_SORT, dispatch.ACC_ESC))
    # Both sparse (inner and outer sizes below 3 * nnz): the lexorder rule;
    # A in [outer inner] and B in [inner outer] -> not transposed, CSC.
    a = _op([20, 30], "ik", 400)
    b = _op([30, 20], "kj", 400)
    assert _heuristic(a, b) == DENSE
    # A in [inner outer] and B in [outer inner] -> transposed.
    a_t = _op([20, 30], "ik", 400, seq=[1, 0])
    b_t = _op([30, 20], "kj", 400, seq=[1, 0])
    assert _heuristic(a_t, b_t) == DENSE | SW
    a3 = _op([20, 30, 5], "ikm", 900, seq=[1, 0, 2])
    b3 = _op([30, 40], "kj", 900, seq=[1, 0])
    assert _heuristic(a3, b3) == DENSE | SW  # a_io and b_oi
    # The fifth case (A in [inner outer], B in neither order): transposed
    # with algorithm 0 (operator*, csc_no_accum).
    a5 = _op([20, 30], "ik", 400, seq=[1, 0])
    b5 = _op([30, 4, 5], "kjm", 400, seq=[1, 0, 2])
    assert _heuristic(a5, b5) == SORT | SW
    # A sparse, B hyper-sparse (inner and outer sizes > 3 * nnz): CSC.
    b_h = _op([1000, 1000], "kj", 50)
    a_s = _op([20, 1000], "ik", 10000)
    assert _heuristic(a_s, b_h) == DENSE
    # A hyper-sparse, B sparse: CSR (transposed).
    a_h = _op([1000, 1000], "ik", 50)
    b_s = _op([1000, 20], "kj", 10000)
    assert _heuristic(a_h, b_s) == DENSE | SW
    # Both hyper-sparse: the outer-product algorithm, not transposed.
    assert _heuristic(a_h, b_h) == ESC
    # A row-sparse only, B col-sparse only, default orders: CSCNA, not transposed.
    assert _heuristic(_op([1000, 10], "ik", 40), _op([10, 1000], "kj", 40)) == SORT
    # A row-sparse only (i: 1000/40 > 3, k: 20/40), B row-sparse only
    # (k: 20/4 > 3, j: 10/4): the DCSR branch, transposed.
    assert _heuristic(_op([1000, 20], "ik", 40), _op([20, 10], "kj", 4)) == DENSE | SW


def test_dispatch_argument_errors():
    import libnt_amd

    a = _op([4, 5], "ij", 3)
    b = _op([5, 6], "jj", 3)  # repeated label
    with pytest.raises(libnt_amd.LntError):
        _heuristic(a, b)
    b = _op([6, 6], "jk", 3)  # j differs in size
    with pytest.raises(libnt_amd.LntError) as e:
        _features(a, b)
    assert e.value.status == 7
    full = _op([4, 5], "ij", 3)
    with pytest.raises(libnt_amd.LntError):  # contracts every mode: a scalar
        _features(full, _op([4, 5], "ij", 3))
