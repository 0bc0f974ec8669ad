"""libnt_amd -- MI355X (gfx950) sparse multi-index-array hot path.

Python host binding of the C-ABI in ``include/lnt.h`` (``liblnt.so``, built by
``libnt_amd/build.py``).  Arrays are torch tensors on the GPU (device memory,
``LNT_MEM_DEVICE``) or numpy arrays (host memory, ``LNT_MEM_HOST``); torch is
only the allocator and stream provider.  Every call runs the HIP kernels; if
the shared library is missing, importing ``lib()`` raises -- there is no CPU
fallback.

``libnt_amd.sparse`` mirrors the reference's ``SparseMIA`` front end on top of
this binding.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblnt.so")

LNT_OK = 0
STATUS_NAMES = {
    0: "LNT_OK", 1: "LNT_ERR_INVALID_ARGUMENT", 2: "LNT_ERR_OUT_OF_MEMORY", 3: "LNT_ERR_DEVICE",
    4: "LNT_ERR_NOT_SORTED", 5: "LNT_ERR_KEY_RANGE", 6: "LNT_ERR_TOO_LARGE",
    7: "LNT_ERR_DIMENSION_MISMATCH", 8: "LNT_ERR_NO_DEVICE",
}
MEM_HOST, MEM_DEVICE = 0, 1
OP_ADD, OP_SUB, OP_HADAMARD = 0, 1, 2
CMP_EXACT, CMP_FUZZY = 0, 1
SPGEMM_AUTO, SPGEMM_SORT, SPGEMM_HASH, SPGEMM_DENSE, SPGEMM_ESC, SPGEMM_MTT = 0, 1, 2, 3, 4, 5
SPGEMM_CSC = 16  # flag: A's columns through a column-pointer array
UNARY_CONTRACT, UNARY_ATTRACT = 0, 1

_lib = None
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64 = ctypes.c_double

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "lnt_abi_version": [],
    "lnt_last_error": [],
    "lnt_context_create": [ctypes.POINTER(_vp), _int, _vp],
    "lnt_context_destroy": [_vp],
    "lnt_context_set_stream": [_vp, _vp],
    "lnt_context_synchronize": [_vp],
    "lnt_context_trim": [_vp],
    "lnt_context_set_validation": [_vp, _int],
    "lnt_context_set_param": [_vp, ctypes.c_char_p, _i64],
    "lnt_context_get_param": [_vp, ctypes.c_char_p, _i64p],
    "lnt_context_set_profiling": [_vp, _int],
    "lnt_context_profile": [_vp, ctypes.c_char_p, _i64p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "lnt_change_linidx": [_vp, _vp, _i64, _int, _i64p, _i32p, _i32p, _int],
    "lnt_sort_pairs": [_vp, _vp, _vp, _i64, _u64, _int],
    "lnt_change_sort_order": [_vp, _vp, _vp, _i64, _int, _i64p, _i32p, _i32p, _int, _int],
    "lnt_merge_into": [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _int, _vp, _vp, _i64, _i64p, _int],
    "lnt_merge_into_noalloc": [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _int, _vp, _vp, _i64, _i64p],
    "lnt_compare": [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _int, ctypes.c_double, ctypes.POINTER(ctypes.c_int), _int],
    "lnt_merge": [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64p, _int],
    "lnt_free": [_vp, _vp, _int],
    "lnt_copy": [_vp, _vp, _vp, _i64],
    "lnt_alloc": [_vp, _i64, _int, ctypes.POINTER(_vp)],
    "lnt_sort_merge_into": [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _int, _u64, _vp, _vp, _i64, _i64p, _int],
    "lnt_sort_merge": [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _int, _u64, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                       _i64p, _int],
    "lnt_lattice_nprod": [_vp, _vp, _i64, _i64p, _vp, _i64, _i64p, _i64p, _int],
    "lnt_lattice_mult_into": [_vp, _vp, _vp, _i64, _i64p, _vp, _vp, _i64, _i64p, _int, _vp, _vp, _i64, _i64p, _int],
    "lnt_lattice_mult": [_vp, _vp, _vp, _i64, _i64p, _vp, _vp, _i64, _i64p, _int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64p, _int],
    "lnt_testgen_distinct_keys": [_vp, _vp, _i64, _u64, _u64, _u64],
    "lnt_testgen_uniform": [_vp, _vp, _i64, _u64, _u64],
    "lnt_testgen_batch": [_vp, _vp, _vp, _i64p, ctypes.POINTER(_u64), _i64, _u64, _u64],
    "lnt_batch_create": [_vp, _vp, _i64, ctypes.POINTER(_vp)],
    "lnt_batch_destroy": [_vp],
    "lnt_batch_contract": [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64p],
    "lnt_batch_symbolic": [_vp, _vp, _vp, _vp, _vp, _i64p],
    "lnt_batch_numeric": [_vp, _vp, _vp, _i64, _i64p],
    "lnt_batch_capture": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64p],
    "lnt_batch_replay": [_vp, _i64p],
    "lnt_batch_launch": [_vp],
    "lnt_batch_set_param": [_vp, ctypes.c_char_p, _i64],
    "lnt_batch_get_param": [_vp, ctypes.c_char_p, _i64p],
    "lnt_batch_set_accumulators": [_vp, _i32p],
    "lnt_batch_nprod": [_vp, _vp, _vp, _i64p],
    "lnt_contract": [_vp, _vp, _vp, _int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64p, _vp, ctypes.POINTER(_int)],
    "lnt_contract_into": [_vp, _vp, _vp, _int, _vp, _vp, _i64, _i64p, _vp, ctypes.POINTER(_int)],
    "lnt_contract_features": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    "lnt_contract_heuristic": [_vp, _vp, ctypes.POINTER(_int)],
    "lnt_contract_time": [_vp, _vp, _vp, _int, _int, ctypes.POINTER(ctypes.c_double)],
    "lnt_model_create": [_vp, _u64, ctypes.POINTER(_vp)],
    "lnt_model_destroy": [_vp],
    "lnt_model_param_count": [],
    "lnt_model_get_params": [_vp, ctypes.POINTER(ctypes.c_float)],
    "lnt_model_set_params": [_vp, ctypes.POINTER(ctypes.c_float)],
    "lnt_model_train": [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _i64, _int, _int,
                        ctypes.c_float, _u64, ctypes.POINTER(ctypes.c_float)],
    "lnt_model_predict": [_vp, ctypes.POINTER(ctypes.c_float), _i64, ctypes.POINTER(ctypes.c_float)],
    "lnt_model_gradient": [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _i64,
                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "lnt_context_set_model": [_vp, _vp],
    "lnt_sparse_dense_mult": [_vp, _vp, _vp, _i64, _i64, _int, _i64p, _vp, _i32p, _int, _i32p, _int,
                              ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64p, _int],
    "lnt_contract_attract": [_vp, _vp, _vp, _i64, _int, _i64p, _i32p, _i32p, _i32p, _int, ctypes.POINTER(_vp),
                             ctypes.POINTER(_vp), _i64p, _int],
    # include/lnt_mex.h (the alloc argument is a C function pointer; None = calloc)
    "lnt_mex_sparse_sort": [_vp, _vp, _vp, _i64],
    "lnt_mex_sparse_change_linidx": [_vp, _vp, _i64, _u8p, _u8p, _i64p, _int],
    "lnt_mex_sparse_permute": [_vp, _vp, _vp, _i64, _u8p, _u8p, _i64p, _int],
    "lnt_mex_sparse_merge": [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _f64, _vp, ctypes.POINTER(_vp),
                             ctypes.POINTER(_vp), _i64p],
    "lnt_mex_sparselattice_mult": [_vp, _vp, _vp, _i64, _i64p, _vp, _vp, _i64, _i64p, _f64, _vp,
                                   ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64p],
}
_RESTYPES = {"lnt_last_error": ctypes.c_char_p, "lnt_model_param_count": ctypes.c_int64}


class LntError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


def lib():
    """Loads liblnt.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libnt_amd/liblnt.so is not built; run `python -m libnt_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != LNT_OK:
        raise LntError(status, lib().lnt_last_error().decode(errors="replace"))


def _arr_i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_i64p)


def _arr_i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def _ptr_mem(x):
    """(pointer, mem kind, length) of a torch GPU tensor or a numpy array."""
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("numpy arrays must be contiguous")
        return x.ctypes.data, MEM_HOST, x.shape[0]
    import torch

    if isinstance(x, torch.Tensor):
        if not x.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return x.data_ptr(), (MEM_DEVICE if x.is_cuda else MEM_HOST), x.shape[0]
    raise TypeError("expected a numpy array or a torch tensor, got %r" % type(x))


class Context:
    """Owns an ``lnt_context`` (device, stream, scratch memory)."""

    def __init__(self, device: int = 0, stream=None, validation: int = 1):
        self._h = _vp()
        check(lib().lnt_context_create(ctypes.byref(self._h), device, stream))
        self.device = device
        if validation != 1:
            self.set_validation(validation)

    def close(self):
        if self._h:
            lib().lnt_context_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream):
        check(lib().lnt_context_set_stream(self._h, stream))

    def synchronize(self):
        check(lib().lnt_context_synchronize(self._h))

    def trim(self):
        check(lib().lnt_context_trim(self._h))

    def set_validation(self, level: int):
        check(lib().lnt_context_set_validation(self._h, level))

    def set_param(self, name: str, value: int):
        check(lib().lnt_context_set_param(self._h, name.encode(), int(value)))

    def get_param(self, name: str) -> int:
        v = ctypes.c_int64()
        check(lib().lnt_context_get_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_profiling(self, on: bool):
        check(lib().lnt_context_set_profiling(self._h, int(bool(on))))

    def profile(self, kernel: str):
        """(launches, total_ms, algorithmic_bytes) recorded for one kernel."""
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        b = ctypes.c_double()
        check(lib().lnt_context_profile(self._h, kernel.encode(), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
        return n.value, ms.value, b.value

    # ------------------------------------------------------------ hot path
    def change_linidx(self, keys, dims, cur_seq, new_seq):
        """In-place relinearisation (SparseMIABase.h:768-814)."""
        kp, mem, n = _ptr_mem(keys)
        d, dp = _arr_i64(dims)
        c, cp = _arr_i32(cur_seq)
        s, sp = _arr_i32(new_seq)
        check(lib().lnt_change_linidx(self._h, kp, n, len(d), dp, cp, sp, mem))

    def sort_pairs(self, keys, vals, max_size):
        """In-place ascending sort of unique keys with values alongside."""
        kp, mem, n = _ptr_mem(keys)
        vp, mem2, n2 = _ptr_mem(vals)
        if mem != mem2 or n != n2:
            raise ValueError("keys and values must match in length and memory kind")
        check(lib().lnt_sort_pairs(self._h, kp, vp, n, int(max_size), mem))

    def change_sort_order(self, keys, vals, dims, cur_seq, new_seq, is_sorted=True):
        """In-place relinearise + sort (SparseMIABase.h:590-645)."""
        kp, mem, n = _ptr_mem(keys)
        vp, mem2, n2 = _ptr_mem(vals)
        if mem != mem2 or n != n2:
            raise ValueError("keys and values must match in length and memory kind")
        d, dp = _arr_i64(dims)
        c, cp = _arr_i32(cur_seq)
        s, sp = _arr_i32(new_seq)
        check(lib().lnt_change_sort_order(self._h, kp, vp, n, len(d), dp, cp, sp, int(bool(is_sorted)), mem))

    def compare(self, ak, av, bk, bv, precision=None) -> bool:
        """compare_with_sparse (SparseMIABase.h:1243-1296) of two operands
        sorted in one linearisation: exact (operator==) when precision is None,
        else isEqualFuzzy(a, b, precision) (fuzzy_equals)."""
        akp, mem, na = _ptr_mem(ak)
        avp, _, _ = _ptr_mem(av)
        bkp, memb, nb = _ptr_mem(bk)
        bvp, _, _ = _ptr_mem(bv)
        if mem != memb:
            raise ValueError("both operands must live in the same memory kind")
        eq = ctypes.c_int(0)
        mode = CMP_EXACT if precision is None else CMP_FUZZY
        check(lib().lnt_compare(self._h, akp, avp, na, bkp, bvp, nb, mode, float(precision or 0.0),
                                ctypes.byref(eq), mem))
        return bool(eq.value)

    def merge_into(self, ak, av, bk, bv, op, ck, cv) -> int:
        """Sorted union (op ADD/SUB) or intersection (HADAMARD) into ck/cv."""
        akp, mem, na = _ptr_mem(ak)
        avp, _, _ = _ptr_mem(av)
        bkp, memb, nb = _ptr_mem(bk)
        bvp, _, _ = _ptr_mem(bv)
        ckp, memc, cap = _ptr_mem(ck)
        cvp, _, capv = _ptr_mem(cv)
        if not (mem == memb == memc):
            raise ValueError("all arrays must live in the same memory kind")
        nc = ctypes.c_int64(0)
        check(lib().lnt_merge_into(self._h, akp, avp, na, bkp, bvp, nb, op, ckp, cvp,
                                   min(cap, capv), ctypes.byref(nc), mem))
        return nc.value

    def merge(self, ak, av, bk, bv, op):
        """Returns new (keys, values) of the merge, of the operands' kind."""
        n_a, n_b = ak.shape[0], bk.shape[0]
        cap = min(n_a, n_b) if op == OP_HADAMARD else n_a + n_b
        if isinstance(ak, np.ndarray):
            ck = np.empty(max(cap, 1), dtype=np.int64)
            cv = np.empty(max(cap, 1), dtype=np.float64)
        else:
            import torch

            ck = torch.empty(max(cap, 1), dtype=torch.int64, device=ak.device)
            cv = torch.empty(max(cap, 1), dtype=torch.float64, device=ak.device)
        n = self.merge_into(ak, av, bk, bv, op, ck, cv)
        return ck[:n], cv[:n]

    def sort_merge_into(self, ak, av, bk, bv, op, bound, ck, cv) -> int:
        """Union of two unsorted operands in one linearisation (SparseMIA::
        sortMerge, with collect_duplicates' zero-entry rule) into ck/cv."""
        akp, mem, na = _ptr_mem(ak)
        avp, _, _ = _ptr_mem(av)
        bkp, memb, nb = _ptr_mem(bk)
        bvp, _, _ = _ptr_mem(bv)
        ckp, memc, cap = _ptr_mem(ck)
        cvp, _, capv = _ptr_mem(cv)
        if not (mem == memb == memc):
            raise ValueError("all arrays must live in the same memory kind")
        nc = ctypes.c_int64(0)
        check(lib().lnt_sort_merge_into(self._h, akp, avp, na, bkp, bvp, nb, op, int(bound), ckp, cvp,
                                        min(cap, capv), ctypes.byref(nc), mem))
        return nc.value

    def sort_merge(self, ak, av, bk, bv, op, bound):
        cap = max(ak.shape[0] + bk.shape[0], 1)
        if isinstance(ak, np.ndarray):
            ck = np.empty(cap, dtype=np.int64)
            cv = np.empty(cap, dtype=np.float64)
        else:
            import torch

            ck = torch.empty(cap, dtype=torch.int64, device=ak.device)
            cv = torch.empty(cap, dtype=torch.float64, device=ak.device)
        n = self.sort_merge_into(ak, av, bk, bv, op, bound, ck, cv)
        return ck[:n], cv[:n]

    def lattice_nprod(self, ak, a_dims, bk, b_dims) -> int:
        """Scalar products of A*B (bound on nnz(C); flops = 2*nprod)."""
        akp, mem, na = _ptr_mem(ak)
        bkp, memb, nb = _ptr_mem(bk)
        ad, adp = _arr_i64(a_dims)
        bd, bdp = _arr_i64(b_dims)
        out = ctypes.c_int64()
        check(lib().lnt_lattice_nprod(self._h, akp, na, adp, bkp, nb, bdp, ctypes.byref(out), mem))
        return out.value

    def lattice_mult_into(self, ak, av, a_dims, bk, bv, b_dims, ck, cv, strategy=SPGEMM_AUTO) -> int:
        akp, mem, na = _ptr_mem(ak)
        avp, _, _ = _ptr_mem(av)
        bkp, memb, nb = _ptr_mem(bk)
        bvp, _, _ = _ptr_mem(bv)
        ckp, memc, cap = _ptr_mem(ck)
        cvp, _, capv = _ptr_mem(cv)
        if not (mem == memb == memc):
            raise ValueError("all arrays must live in the same memory kind")
        ad, adp = _arr_i64(a_dims)
        bd, bdp = _arr_i64(b_dims)
        nc = ctypes.c_int64()
        check(lib().lnt_lattice_mult_into(self._h, akp, avp, na, adp, bkp, bvp, nb, bdp, strategy,
                                          ckp, cvp, min(cap, capv), ctypes.byref(nc), mem))
        return nc.value

    def lattice_mult(self, ak, av, a_dims, bk, bv, b_dims, strategy=SPGEMM_AUTO):
        """C = A*B per tab; returns (keys, values) of C's kind, sorted."""
        cap = max(self.lattice_nprod(ak, a_dims, bk, b_dims), 1)
        if isinstance(ak, np.ndarray):
            ck = np.empty(cap, dtype=np.int64)
            cv = np.empty(cap, dtype=np.float64)
        else:
            import torch

            ck = torch.empty(cap, dtype=torch.int64, device=ak.device)
            cv = torch.empty(cap, dtype=torch.float64, device=ak.device)
        n = self.lattice_mult_into(ak, av, a_dims, bk, bv, b_dims, ck, cv, strategy)
        return ck[:n], cv[:n]

    def _take_result(self, ck, cv, m, mem, like):
        """Copies a library-allocated result of m pairs into arrays of the
        operand's kind and frees it."""
        try:
            if mem == MEM_HOST:
                ok = np.ctypeslib.as_array(ctypes.cast(ck, _i64p), (max(m, 1),))[:m].copy()
                ov = np.ctypeslib.as_array(ctypes.cast(cv, ctypes.POINTER(ctypes.c_double)), (max(m, 1),))[:m].copy()
                return ok, ov
            import torch

            ok = torch.empty(m, dtype=torch.int64, device=like.device)
            ov = torch.empty(m, dtype=torch.float64, device=like.device)
            if m:
                check(lib().lnt_copy(self._h, ok.data_ptr(), ck, m * 8))
                check(lib().lnt_copy(self._h, ov.data_ptr(), cv, m * 8))
            return ok, ov
        finally:
            lib().lnt_free(self._h, ck, mem)
            lib().lnt_free(self._h, cv, mem)

    def sparse_dense_mult(self, ak, av, l_size, b_flat, b_dims, r_modes, i_modes):
        """lnt_sparse_dense_mult: A sorted with key = l + l_size*t; b_flat the
        dense B's elements, first mode fastest (a 1-D tensor or numpy array of
        A's kind), with dims b_dims, modes r_modes outer and i_modes
        element-wise.  Returns C's (keys, values), key r + l*R + t*l_size*R."""
        akp, mem, n = _ptr_mem(ak)
        avp, _, _ = _ptr_mem(av)
        bp, bmem, nb = _ptr_mem(b_flat)
        if bmem != mem:
            raise ValueError("A and B must live in the same memory kind")
        if nb != int(np.prod(b_dims)):
            raise ValueError("b_flat must hold prod(b_dims) elements")
        d, dp = _arr_i64(b_dims)
        rm, rp = _arr_i32(list(r_modes) or [0])
        im, ip = _arr_i32(list(i_modes) or [0])
        ck, cv, nc = _vp(), _vp(), ctypes.c_int64()
        check(lib().lnt_sparse_dense_mult(self._h, akp, avp, n, int(l_size), len(d), dp, bp, rp, len(r_modes), ip,
                                          len(i_modes), ctypes.byref(ck), ctypes.byref(cv), ctypes.byref(nc), mem))
        return self._take_result(ck, cv, nc.value, mem, ak)

    def contract_attract(self, keys, vals, dims, seq, group, kind):
        """Unary contraction/attraction (lnt_contract_attract): group[m] 0 keeps
        mode m, g >= 1 puts it in group g; kind[g-1] UNARY_CONTRACT or
        UNARY_ATTRACT.  Returns (keys, values) of the result, of the operand's
        kind, sorted, in the result's default order (staying modes, then the
        attraction groups)."""
        kp, mem, n = _ptr_mem(keys)
        vp, _, _ = _ptr_mem(vals)
        d, dp = _arr_i64(dims)
        sq, sp = _arr_i32(seq)
        gr, gp = _arr_i32(group)
        kd, kdp = _arr_i32(kind if len(kind) else [0])
        ck, cv, nc = _vp(), _vp(), ctypes.c_int64()
        check(lib().lnt_contract_attract(self._h, kp, vp, n, len(d), dp, sp, gp, kdp, len(kind), ctypes.byref(ck),
                                         ctypes.byref(cv), ctypes.byref(nc), mem))
        return self._take_result(ck, cv, nc.value, mem, keys)

    # ---------------------------------------------------- synthetic inputs
    def testgen_distinct_keys(self, keys, domain, seed, offset=0):
        kp, mem, n = _ptr_mem(keys)
        if mem != MEM_DEVICE:
            raise ValueError("testgen writes device memory")
        check(lib().lnt_testgen_distinct_keys(self._h, kp, n, domain, seed, offset))

    def testgen_uniform(self, vals, seed, offset=0):
        vp, mem, n = _ptr_mem(vals)
        if mem != MEM_DEVICE:
            raise ValueError("testgen writes device memory")
        check(lib().lnt_testgen_uniform(self._h, vp, n, seed, offset))

    def testgen_batch(self, keys, vals, offsets, domains, key_seed, val_seed):
        """Per-instance keys/values (instance i: seeds key_seed+i, val_seed+i)."""
        kp = _ptr_mem(keys)[0] if keys is not None else None
        vp = _ptr_mem(vals)[0] if vals is not None else None
        o, op = _arr_i64(offsets)
        d = np.ascontiguousarray(domains, dtype=np.uint64)
        check(lib().lnt_testgen_batch(self._h, kp, vp, op, d.ctypes.data_as(ctypes.POINTER(_u64)),
                                      len(d), key_seed, val_seed))


def exported_symbols_declared_in_headers(include_dir=None):
    """Names of the functions declared in include/*.h (used by the ABI test)."""
    import glob
    import re

    include_dir = include_dir or os.path.join(os.path.dirname(HERE), "include")
    names = set()
    for h in glob.glob(os.path.join(include_dir, "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(lnt_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)
