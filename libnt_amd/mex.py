"""The reference's MATLAB mex entry points on numpy arrays
(``include/lnt_mex.h``): the same argument order and meaning as the mex calls
in the reference's NTToolbox (``src/MATLAB/mex/*.cpp``) -- 1-based uint8 mode
sequences, int64 0-based indices, the in-place entries mutating the caller's
arrays, the others returning new arrays.  Every call runs the library's HIP
kernels through the C-ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _i64p, _u8p, _vp, check, lib

_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]
_libc.free.restype = None


def _vec(a, dtype, name):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or a.ndim != 1 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("%s must be a contiguous 1-D %s numpy array" % (name, np.dtype(dtype).name))
    return a


def _pairs(data, idx):
    _vec(data, np.float64, "data")
    _vec(idx, np.int64, "idx")
    if data.shape != idx.shape:
        raise ValueError("Data and index vectors must be the same length.")
    return data.shape[0]


def _seqs(cur_seq, new_seq, dims):
    cur = np.ascontiguousarray(cur_seq, dtype=np.uint8)
    new = np.ascontiguousarray(new_seq, dtype=np.uint8)
    d = np.ascontiguousarray(dims, dtype=np.int64)
    if not (cur.shape == new.shape == d.shape) or cur.ndim != 1:
        raise ValueError("Lexicographical indices vectors and dimensions vectors must be same size.")
    return cur, new, d


def _take(kp, vp, n):
    """Copies a calloc'd result into numpy arrays and frees it."""
    try:
        if n == 0:
            return np.empty(0, np.float64), np.empty(0, np.int64)
        v = np.ctypeslib.as_array(ctypes.cast(vp, ctypes.POINTER(ctypes.c_double)), (n,)).copy()
        k = np.ctypeslib.as_array(ctypes.cast(kp, ctypes.POINTER(ctypes.c_int64)), (n,)).copy()
        return v, k
    finally:
        _libc.free(vp)
        _libc.free(kp)


def sparse_sort(ctx, data: np.ndarray, idx: np.ndarray) -> None:
    """SparseSortMex(data, idx): sorts idx ascending in place, data alongside."""
    n = _pairs(data, idx)
    check(lib().lnt_mex_sparse_sort(ctx._h, data.ctypes.data, idx.ctypes.data, n))


def sparse_change_linidx(ctx, idx: np.ndarray, cur_seq, new_seq, dims) -> None:
    """SparseChangeLinIdxMex(idx, curSeq, newSeq, dims), in place; 1-based sequences."""
    _vec(idx, np.int64, "idx")
    cur, new, d = _seqs(cur_seq, new_seq, dims)
    check(lib().lnt_mex_sparse_change_linidx(ctx._h, idx.ctypes.data, idx.shape[0], cur.ctypes.data_as(_u8p),
                                             new.ctypes.data_as(_u8p), d.ctypes.data_as(_i64p), d.shape[0]))


def sparse_permute(ctx, data: np.ndarray, idx: np.ndarray, cur_seq, new_seq, dims) -> None:
    """SparsePermuteMex(data, idx, curSeq, newSeq, dims), in place: idx already
    in the new_seq linearisation, ordered as a tensor sorted in cur_seq."""
    n = _pairs(data, idx)
    cur, new, d = _seqs(cur_seq, new_seq, dims)
    check(lib().lnt_mex_sparse_permute(ctx._h, data.ctypes.data, idx.ctypes.data, n, cur.ctypes.data_as(_u8p),
                                       new.ctypes.data_as(_u8p), d.ctypes.data_as(_i64p), d.shape[0]))


def sparse_merge(ctx, a_data, a_idx, b_data, b_idx, op: float):
    """[c_data, c_idx] = SparseMergeMex(a_data, a_idx, b_data, b_idx, op)."""
    na, nb = _pairs(a_data, a_idx), _pairs(b_data, b_idx)
    kp, vp, nc = _vp(), _vp(), ctypes.c_int64()
    check(lib().lnt_mex_sparse_merge(ctx._h, a_data.ctypes.data, a_idx.ctypes.data, na, b_data.ctypes.data,
                                     b_idx.ctypes.data, nb, float(op), None, ctypes.byref(vp), ctypes.byref(kp),
                                     ctypes.byref(nc)))
    return _take(kp, vp, nc.value)


def sparselattice_mult(ctx, a_data, a_idx, a_dims, b_data, b_idx, b_dims, algorithm: float = 0):
    """[c_data, c_idx] = SparseLatticeMultMex(a_data, a_idx, [h w d], b_data,
    b_idx, [h w d], algorithm)."""
    na, nb = _pairs(a_data, a_idx), _pairs(b_data, b_idx)
    ad = np.ascontiguousarray(a_dims, dtype=np.int64)
    bd = np.ascontiguousarray(b_dims, dtype=np.int64)
    if ad.shape != (3,) or bd.shape != (3,):
        raise ValueError("Must input height, width, and depth dimensions for each input.")
    kp, vp, nc = _vp(), _vp(), ctypes.c_int64()
    check(lib().lnt_mex_sparselattice_mult(ctx._h, a_data.ctypes.data, a_idx.ctypes.data, na,
                                           ad.ctypes.data_as(_i64p), b_data.ctypes.data, b_idx.ctypes.data, nb,
                                           bd.ctypes.data_as(_i64p), float(algorithm), None, ctypes.byref(vp),
                                           ctypes.byref(kp), ctypes.byref(nc)))
    return _take(kp, vp, nc.value)
