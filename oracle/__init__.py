"""CPU oracle for the sparse multi-index-array hot path (test infrastructure).

A ctypes wrapper over ``liblnt_oracle.so`` (built from ``lnt_oracle.c`` by
``oracle/Makefile``).  This package is the CHECKER: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product library (``libnt_amd``) never imports or calls it.

All arrays are numpy: keys ``int64``, values ``float64``, dims ``int64``,
mode sequences ``int32`` (first entry = fastest-varying mode, as in the
reference's ``linIdxSequence``, SparseMIABase.h:98-153).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblnt_oracle.so")
_lib = None

OP_ADD, OP_SUB, OP_HADAMARD = 0, 1, 2

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)


def _stale() -> bool:
    t = os.path.getmtime(_LIB_PATH)
    return any(os.path.getmtime(os.path.join(_HERE, f)) > t for f in ("lnt_oracle.c", "lnt_oracle.h", "Makefile"))


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no GPU, no reference sources needed); make
    rebuilds it when its sources are newer than the library."""
    if force or not os.path.exists(_LIB_PATH) or _stale():
        subprocess.check_call(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, i64, f64, i32 = ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        sig = {
            "lnto_mix64": (u64, [u64]),
            "lnto_collect_duplicates_member": (i64, [_i64p, _f64p, i64]),
            "lnto_sort_merge": (i64, [_i64p, _f64p, i64, _i64p, _f64p, i64, i32, u64, _i64p, _f64p]),
            "lnto_mlp_param_count": (i64, []),
            "lnto_mlp_weight_count": (i64, []),
            "lnto_mlp_forward": (None, [_f64p, _f64p, _f64p]),
            "lnto_mlp_loss_grad": (f64, [_f64p, _f64p, _f64p, i64, _f64p]),
            "lnto_rand_u64": (u64, [u64, u64]),
            "lnto_uniform_pm1": (f64, [u64, u64]),
            "lnto_uniform_01": (f64, [u64, u64]),
            "lnto_normal": (f64, [u64, u64]),
            "lnto_feistel": (u64, [u64, i32, u64]),
            "lnto_permute_index": (u64, [u64, u64, u64]),
            "lnto_gen_distinct_keys": (None, [_i64p, i64, u64, u64, u64]),
            "lnto_gen_uniform_vals": (None, [_f64p, i64, u64, u64]),
            "lnto_random_perm": (None, [_i32p, i32, u64]),
            "lnto_gen_coords": (i32, [_i64p, _f64p, i64, i32, _i64p, _i32p, f64, i32, u64]),
            "lnto_sub2ind": (i64, [_i64p, _i32p, _i64p, i32]),
            "lnto_ind2sub": (None, [i64, _i32p, _i64p, i32, _i64p]),
            "lnto_change_linidx": (None, [_i64p, i64, i32, _i64p, _i32p, _i32p]),
            "lnto_radix_sort_pairs": (None, [_i64p, _f64p, i64, u64]),
            "lnto_introsort_pairs": (None, [_i64p, _f64p, i64]),
            "lnto_sort_seq": (None, [_i64p, _f64p, i64, i32, _i64p, _i32p, _i32p, i32]),
            "lnto_sort_seq_shuffle": (None, [_i64p, _f64p, i64, i32, _i64p, _i32p, _i32p, i32]),
            "lnto_collect_duplicates": (i64, [_i64p, _f64p, i64]),
            "lnto_merge": (i64, [_i64p, _f64p, i64, _i64p, _f64p, i64, i32, _i64p, _f64p]),
            "lnto_mex_merge": (i64, [_i64p, _f64p, i64, _i64p, _f64p, i64, i32, _i64p, _f64p]),
            "lnto_outer_inter_count": (i64, [_i64p, i64, i64, _i64p, i64, i64]),
            "lnto_outer_inter_product": (i64, [_i64p, _f64p, i64, i64, _i64p, _f64p, i64, i64, _i64p, _f64p]),
            "lnto_lattice_nprod": (i64, [_i64p, i64, i64, i64, _i64p, i64, i64, i64]),
            "lnto_lattice_mult": (i64, [_i64p, _f64p, i64, i64, i64, i64, _i64p, _f64p, i64, i64, i64, i64, _i64p, _f64p, i64]),
            "lnto_lattice_mult_spa": (i64, [_i64p, _f64p, i64, i64, i64, _i64p, _f64p, i64, i64, i64, i32, _i64p,
                                            _f64p, i64]),
            "lnto_lattice_mult_outer": (i64, [_i64p, _f64p, i64, i64, i64, _i64p, _f64p, i64, i64, i64, _i64p,
                                              _f64p, i64]),
            "lnto_contract_plan": (i32, [i32, _i64p, ctypes.c_char_p, i32, _i64p, ctypes.c_char_p, _i32p, _i32p, _i64p, _i64p, _i64p, ctypes.POINTER(ctypes.c_int)]),
            "lnto_is_equal_fuzzy": (i32, [f64, f64, f64]),
            "lnto_compare_sparse": (i32, [_i64p, _f64p, i64, _i64p, _f64p, i64, i32, f64]),
            "lnto_sparse_dense_mult": (i64, [_i64p, _f64p, i64, i64, i32, _i64p, _f64p, _i32p, i32, _i32p, i32,
                                             _i64p, _f64p]),
            "lnto_contract_attract": (i64, [_i64p, _f64p, i64, i32, _i64p, _i32p, _i32p, _i32p, i32, _i64p, _f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------- generators
def mix64(x: int) -> int:
    return lib().lnto_mix64(x)


def rand_u64(seed: int, i: int) -> int:
    return lib().lnto_rand_u64(seed, i)


def uniform_pm1(seed: int, i: int) -> float:
    return lib().lnto_uniform_pm1(seed, i)


def feistel(x: int, bits: int, seed: int) -> int:
    return lib().lnto_feistel(x, bits, seed)


def permute_index(i: int, domain: int, seed: int) -> int:
    return lib().lnto_permute_index(i, domain, seed)


def gen_distinct_keys(n: int, domain: int, seed: int, offset: int = 0) -> np.ndarray:
    k = np.empty(n, dtype=np.int64)
    lib().lnto_gen_distinct_keys(_p(k, _i64p), n, domain, seed, offset)
    return k


def gen_uniform_vals(n: int, seed: int, offset: int = 0) -> np.ndarray:
    v = np.empty(n, dtype=np.float64)
    lib().lnto_gen_uniform_vals(_p(v, _f64p), n, seed, offset)
    return v


def random_perm(n: int, seed: int) -> np.ndarray:
    p = np.empty(n, dtype=np.int32)
    lib().lnto_random_perm(_p(p, _i32p), n, seed)
    return p


def gen_coords(nnz, dims, dist, zipf_s=1.1, value_kind=1, seed=0):
    dims = _i64(dims)
    dist = _i32(dist)
    k = np.empty(nnz, dtype=np.int64)
    v = np.empty(nnz, dtype=np.float64)
    rc = lib().lnto_gen_coords(_p(k, _i64p), _p(v, _f64p), nnz, len(dims), _p(dims, _i64p),
                               _p(dist, _i32p), zipf_s, value_kind, seed)
    if rc != 0:
        raise ValueError("gen_coords: cannot draw %d distinct coordinates" % nnz)
    return k, v


# ----------------------------------------------------------------- index math
def change_linidx(keys, dims, cur_seq, new_seq) -> np.ndarray:
    k = _i64(keys).copy()
    dims, cur, new = _i64(dims), _i32(cur_seq), _i32(new_seq)
    lib().lnto_change_linidx(_p(k, _i64p), len(k), len(dims), _p(dims, _i64p), _p(cur, _i32p), _p(new, _i32p))
    return k


def ind2sub(key: int, dims, seq):
    dims, seq = _i64(dims), _i32(seq)
    sub = np.empty(len(dims), dtype=np.int64)
    lib().lnto_ind2sub(key, _p(seq, _i32p), _p(dims, _i64p), len(dims), _p(sub, _i64p))
    return sub


def sub2ind(sub, dims, seq) -> int:
    sub, dims, seq = _i64(sub), _i64(dims), _i32(seq)
    return lib().lnto_sub2ind(_p(sub, _i64p), _p(seq, _i32p), _p(dims, _i64p), len(dims))


# -------------------------------------------------------------------- sorting
def radix_sort_pairs(keys, vals, max_size):
    k, v = _i64(keys).copy(), _f64(vals).copy()
    lib().lnto_radix_sort_pairs(_p(k, _i64p), _p(v, _f64p), len(k), int(max_size))
    return k, v


def introsort_pairs(keys, vals):
    k, v = _i64(keys).copy(), _f64(vals).copy()
    lib().lnto_introsort_pairs(_p(k, _i64p), _p(v, _f64p), len(k))
    return k, v


def sort_seq(keys, vals, dims, cur_seq, new_seq, is_sorted):
    k, v = _i64(keys).copy(), _f64(vals).copy()
    dims, cur, new = _i64(dims), _i32(cur_seq), _i32(new_seq)
    lib().lnto_sort_seq(_p(k, _i64p), _p(v, _f64p), len(k), len(dims), _p(dims, _i64p),
                        _p(cur, _i32p), _p(new, _i32p), int(bool(is_sorted)))
    return k, v


def sort_seq_shuffle(keys, vals, dims, cur_seq, new_seq, is_sorted):
    """sort(seq) by the reference's own RadixShuffle path (sorted input)."""
    k, v = _i64(keys).copy(), _f64(vals).copy()
    sort_seq_shuffle_inplace(k, v, dims, cur_seq, new_seq, is_sorted)
    return k, v


def sort_seq_shuffle_inplace(k, v, dims, cur_seq, new_seq, is_sorted):
    dims, cur, new = _i64(dims), _i32(cur_seq), _i32(new_seq)
    lib().lnto_sort_seq_shuffle(_p(k, _i64p), _p(v, _f64p), len(k), len(dims), _p(dims, _i64p),
                                _p(cur, _i32p), _p(new, _i32p), int(bool(is_sorted)))


def sort_seq_inplace(k, v, dims, cur_seq, new_seq, is_sorted):
    dims, cur, new = _i64(dims), _i32(cur_seq), _i32(new_seq)
    lib().lnto_sort_seq(_p(k, _i64p), _p(v, _f64p), len(k), len(dims), _p(dims, _i64p),
                        _p(cur, _i32p), _p(new, _i32p), int(bool(is_sorted)))


def collect_duplicates(keys, vals):
    k, v = _i64(keys).copy(), _f64(vals).copy()
    n = lib().lnto_collect_duplicates(_p(k, _i64p), _p(v, _f64p), len(k))
    return k[:n].copy(), v[:n].copy()


# -------------------------------------------------------------------- merging
def merge(ak, av, bk, bv, op):
    ak, av, bk, bv = _i64(ak), _f64(av), _i64(bk), _f64(bv)
    cap = len(ak) + len(bk) if op != OP_HADAMARD else min(len(ak), len(bk))
    ck = np.empty(max(cap, 1), dtype=np.int64)
    cv = np.empty(max(cap, 1), dtype=np.float64)
    n = lib().lnto_merge(_p(ak, _i64p), _p(av, _f64p), len(ak), _p(bk, _i64p), _p(bv, _f64p), len(bk),
                         op, _p(ck, _i64p), _p(cv, _f64p))
    return ck[:n].copy(), cv[:n].copy()


def mex_merge(ak, av, bk, bv, op):
    """SparseMergeMex (op 0 add, 1 subtract), its empty-operand cases included."""
    ak, av, bk, bv = _i64(ak), _f64(av), _i64(bk), _f64(bv)
    cap = max(len(ak) + len(bk), 1)
    ck = np.empty(cap, dtype=np.int64)
    cv = np.empty(cap, dtype=np.float64)
    n = lib().lnto_mex_merge(_p(ak, _i64p), _p(av, _f64p), len(ak), _p(bk, _i64p), _p(bv, _f64p), len(bk),
                             op, _p(ck, _i64p), _p(cv, _f64p))
    return ck[:n].copy(), cv[:n].copy()


def outer_inter_product(ak, av, l_outer, bk, bv, r_outer):
    ak, av, bk, bv = _i64(ak), _f64(av), _i64(bk), _f64(bv)
    n = lib().lnto_outer_inter_count(_p(ak, _i64p), len(ak), l_outer, _p(bk, _i64p), len(bk), r_outer)
    ck = np.empty(max(n, 1), dtype=np.int64)
    cv = np.empty(max(n, 1), dtype=np.float64)
    lib().lnto_outer_inter_product(_p(ak, _i64p), _p(av, _f64p), len(ak), l_outer, _p(bk, _i64p),
                                   _p(bv, _f64p), len(bk), r_outer, _p(ck, _i64p), _p(cv, _f64p))
    return ck[:n].copy(), cv[:n].copy()


# ------------------------------------------------------------ lattice product
def lattice_nprod(ak, a_dims3, bk, b_dims3) -> int:
    ak, bk = _i64(ak), _i64(bk)
    return lib().lnto_lattice_nprod(_p(ak, _i64p), len(ak), a_dims3[0], a_dims3[1],
                                    _p(bk, _i64p), len(bk), b_dims3[0], b_dims3[1])


def lattice_mult(ak, av, a_dims3, bk, bv, b_dims3):
    """Both operands must be sorted (column major lattice keys)."""
    ak, av, bk, bv = _i64(ak), _f64(av), _i64(bk), _f64(bv)
    cap = max(lattice_nprod(ak, a_dims3, bk, b_dims3), 1)
    ck = np.empty(cap, dtype=np.int64)
    cv = np.empty(cap, dtype=np.float64)
    n = lib().lnto_lattice_mult(_p(ak, _i64p), _p(av, _f64p), len(ak), *a_dims3,
                                _p(bk, _i64p), _p(bv, _f64p), len(bk), *b_dims3,
                                _p(ck, _i64p), _p(cv, _f64p), cap)
    if n < 0:
        raise ValueError("lattice_mult: dimension mismatch")
    return ck[:n].copy(), cv[:n].copy()


def lattice_mult_by(method, ak, av, a_dims3, bk, bv, b_dims3):
    """The lattice product by one of the reference's other methods: "spa"
    (csc_times), "mtt" (mtt_times, marker zeroed per tab) or "outer"
    (outer_times).  Both operands sorted (column major lattice keys)."""
    ak, av, bk, bv = _i64(ak), _f64(av), _i64(bk), _f64(bv)
    cap = max(lattice_nprod(ak, a_dims3, bk, b_dims3), 1)
    ck = np.empty(cap, dtype=np.int64)
    cv = np.empty(cap, dtype=np.float64)
    args = (_p(ak, _i64p), _p(av, _f64p), len(ak), a_dims3[0], a_dims3[1], _p(bk, _i64p), _p(bv, _f64p), len(bk),
            b_dims3[0], b_dims3[1])
    if method in ("spa", "mtt"):
        n = lib().lnto_lattice_mult_spa(*args, int(method == "mtt"), _p(ck, _i64p), _p(cv, _f64p), cap)
    elif method == "outer":
        n = lib().lnto_lattice_mult_outer(*args, _p(ck, _i64p), _p(cv, _f64p), cap)
    else:
        raise ValueError("unknown method %r" % method)
    if n < 0:
        raise ValueError("lattice_mult_by: dimension mismatch")
    return ck[:n].copy(), cv[:n].copy()


def contract_plan(dims_a, labels_a, dims_b, labels_b):
    """Returns (seq_a, seq_b, lat_a, lat_b, dims_c) for A(labels_a)*B(labels_b)."""
    dims_a, dims_b = _i64(dims_a), _i64(dims_b)
    seq_a = np.empty(len(dims_a), dtype=np.int32)
    seq_b = np.empty(len(dims_b), dtype=np.int32)
    lat_a = np.empty(3, dtype=np.int64)
    lat_b = np.empty(3, dtype=np.int64)
    dims_c = np.empty(len(dims_a) + len(dims_b), dtype=np.int64)
    order_c = ctypes.c_int(0)
    rc = lib().lnto_contract_plan(len(dims_a), _p(dims_a, _i64p), labels_a.encode(), len(dims_b),
                                  _p(dims_b, _i64p), labels_b.encode(), _p(seq_a, _i32p), _p(seq_b, _i32p),
                                  _p(lat_a, _i64p), _p(lat_b, _i64p), _p(dims_c, _i64p), ctypes.byref(order_c))
    if rc != 0:
        raise ValueError("contract_plan: invalid labels %r * %r" % (labels_a, labels_b))
    return seq_a, seq_b, lat_a, lat_b, dims_c[: order_c.value].copy()


def contract(ak, av, dims_a, seq_a_cur, a_sorted, labels_a, bk, bv, dims_b, seq_b_cur, b_sorted, labels_b,
             method="csc_no_accum"):
    """Full MIA contraction; returns (ck, cv, dims_c).  Inputs are not modified.
    method: the lattice product -- "csc_no_accum" (operator*), or "spa",
    "mtt", "outer" (lattice_mult_by)."""
    seq_a, seq_b, lat_a, lat_b, dims_c = contract_plan(dims_a, labels_a, dims_b, labels_b)
    ak2, av2 = sort_seq(ak, av, dims_a, seq_a_cur, seq_a, a_sorted)
    bk2, bv2 = sort_seq(bk, bv, dims_b, seq_b_cur, seq_b, b_sorted)
    la, lb = tuple(int(x) for x in lat_a), tuple(int(x) for x in lat_b)
    if method == "csc_no_accum":
        ck, cv = lattice_mult(ak2, av2, la, bk2, bv2, lb)
    else:
        ck, cv = lattice_mult_by(method, ak2, av2, la, bk2, bv2, lb)
    return ck, cv, dims_c


def sparse_dense_mult(ak, av, l_size, b, b_dims, r_modes, i_modes):
    """noLatticeMult with a dense B (SparseMIABase.h:1829-1910).  A sorted with
    key = l + L*t; b dense (first mode fastest); returns (keys, values) in
    C's key r + l*R + t*L*R, ascending."""
    ak, av = _i64(ak), _f64(av)
    bd, b = _i64(b_dims), _f64(np.ravel(b, order="F"))
    rm, im = _i32(r_modes), _i32(i_modes)
    R = int(np.prod([b_dims[m] for m in r_modes])) if len(r_modes) else 1
    cap = max(len(ak) * R, 1)
    ck = np.empty(cap, dtype=np.int64)
    cv = np.empty(cap, dtype=np.float64)
    n = lib().lnto_sparse_dense_mult(_p(ak, _i64p), _p(av, _f64p), len(ak), int(l_size), len(bd), _p(bd, _i64p),
                                     _p(b, _f64p), _p(rm, _i32p), len(rm), _p(im, _i32p), len(im), _p(ck, _i64p),
                                     _p(cv, _f64p))
    return ck[:n].copy(), cv[:n].copy()


def contract_attract(keys, vals, dims, seq, group, kind):
    """SparseMIABase::contract_attract (SparseMIABase.h:1913-2053): group[m] 0
    keeps mode m, g >= 1 puts it in group g; kind[g-1] 0 contracts, 1
    attracts.  Returns (keys, values) of the result, sorted by key."""
    k, v = _i64(keys).copy(), _f64(vals).copy()
    dims, seq, group, kind = _i64(dims), _i32(seq), _i32(group), _i32(kind)
    ck = np.empty(max(len(k), 1), dtype=np.int64)
    cv = np.empty(max(len(k), 1), dtype=np.float64)
    n = lib().lnto_contract_attract(_p(k, _i64p), _p(v, _f64p), len(k), len(dims), _p(dims, _i64p), _p(seq, _i32p),
                                    _p(group, _i32p), _p(kind, _i32p), len(kind), _p(ck, _i64p), _p(cv, _f64p))
    return ck[:n].copy(), cv[:n].copy()


def is_equal_fuzzy(a: float, b: float, prec: float) -> bool:
    return bool(lib().lnto_is_equal_fuzzy(a, b, prec))


def compare_sparse(ak, av, bk, bv, mode=0, prec=1e-6) -> bool:
    """compare_with_sparse (SparseMIABase.h:1243-1296) of two operands sorted in
    one linearisation; mode 0 exact, 1 isEqualFuzzy(prec)."""
    ak, av, bk, bv = _i64(ak), _f64(av), _i64(bk), _f64(bv)
    return bool(lib().lnto_compare_sparse(_p(ak, _i64p), _p(av, _f64p), len(ak), _p(bk, _i64p), _p(bv, _f64p), len(bk),
                                          int(mode), float(prec)))


def allclose_fuzzy(a, b, prec=1e-12) -> bool:
    """Vectorised FunctionUtil.h:688-698 isEqualFuzzy."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    absd = np.abs(a - b)
    small = (np.abs(a) < 1) | (np.abs(b) < 1)
    ok = np.where(small, absd <= prec, absd <= np.minimum(np.abs(a), np.abs(b)) * prec)
    return bool(np.all(ok))


# ------------------------------------------------------------------ cost model
MLP_OUTPUTS = 20  # LNTO_MLP_S: one predicted log2 time per strategy


def mlp_param_count() -> int:
    return int(lib().lnto_mlp_param_count())


def mlp_weight_count() -> int:
    return int(lib().lnto_mlp_weight_count())


def mlp_forward(params, X) -> np.ndarray:
    """Predictions (n x MLP_OUTPUTS) of the dispatch model for features X (n x 18)."""
    P = _f64(params)
    X = np.ascontiguousarray(X, dtype=np.float64)
    out = np.empty((X.shape[0], MLP_OUTPUTS))
    for i in range(X.shape[0]):
        lib().lnto_mlp_forward(_p(P, _f64p), _p(X[i], _f64p), _p(out[i], _f64p))
    return out


def mlp_loss_grad(params, X, T):
    """(loss, gradient over the trainable parameters) in FP64."""
    P = _f64(params)
    X = np.ascontiguousarray(X, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    g = np.zeros(mlp_weight_count())
    loss = lib().lnto_mlp_loss_grad(_p(P, _f64p), _p(X, _f64p), _p(T, _f64p), X.shape[0], _p(g, _f64p))
    return float(loss), g


def collect_duplicates_member(keys, vals):
    """SparseMIA::collect_duplicates with plus on sorted keys (zero first entries dropped)."""
    k, v = _i64(keys).copy(), _f64(vals).copy()
    n = lib().lnto_collect_duplicates_member(_p(k, _i64p), _p(v, _f64p), len(k))
    return k[:n], v[:n]


def sort_merge(ak, av, bk, bv, op, bound):
    """SparseMIA::sortMerge of operands in one linearisation; returns (keys, values)."""
    ak, av, bk, bv = _i64(ak), _f64(av), _i64(bk), _f64(bv)
    ck = np.empty(len(ak) + len(bk), dtype=np.int64)
    cv = np.empty(len(ak) + len(bk))
    n = lib().lnto_sort_merge(_p(ak, _i64p), _p(av, _f64p), len(ak), _p(bk, _i64p), _p(bv, _f64p), len(bk), op,
                              int(bound), _p(ck, _i64p), _p(cv, _f64p))
    return ck[:n], cv[:n]
