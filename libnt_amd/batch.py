"""Batched contractions (``include/lnt_batch.h``): many independent
``C = A * B`` in one pipeline, optionally recorded as a hipGraph.

The reference evaluates one expression at a time (``MIA_Expr.h:339-419``); a
batch runs the same unfold -> lattice product -> refold for every instance with
all instances sharing each kernel launch.  ``lattice_plan`` is the unfold
bookkeeping of ``MIA_Expr.h:449-561`` (also used by ``libnt_amd.sparse``).
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import MEM_DEVICE, _arr_i64, _i64p, _ptr_mem, _vp, check, lib

MAX_ORDER = 16


def _prod(xs):
    p = 1
    for x in xs:
        p *= int(x)
    return p


def lattice_plan(dims_a: Sequence[int], la: str, dims_b: Sequence[int], lb: str):
    """Unfold of ``C = A(la) * B(lb)``.  Lower-case labels shared by A and B are
    contracted, upper-case shared labels are inter (Hadamard-like) indices,
    unshared labels are outer.  A's lattice: rows = A-outer, columns =
    contracted, tabs = inter, each in A's mode order; B's: rows = contracted in
    A's order, columns = B-outer, tabs = inter in A's order.  C's modes are
    [A-outer, B-outer, inter].

    Returns (seq_a, seq_b, lat_a, lat_b, dims_c, labels_c, (a_in, a_out, b_out))."""
    if len(set(la)) != len(la) or len(set(lb)) != len(lb):
        raise ValueError("an index label repeats within one operand")
    a_out, a_in, a_int, b_in, b_int = [], [], [], [], []
    for m, c in enumerate(la):
        if c in lb:
            r = lb.index(c)
            if int(dims_a[m]) != int(dims_b[r]):
                raise ValueError("matched indices differ in size")
            if c.isupper():
                a_int.append(m)
                b_int.append(r)
            else:
                a_in.append(m)
                b_in.append(r)
        else:
            if c.isupper():
                raise ValueError("inter index %r appears in one operand only" % c)
            a_out.append(m)
    b_out = [r for r, c in enumerate(lb) if c not in la]
    if any(lb[r].isupper() for r in b_out):
        raise ValueError("inter index appears in one operand only")
    seq_a = a_out + a_in + a_int
    seq_b = b_in + b_out + b_int
    lat_a = (_prod(dims_a[m] for m in a_out), _prod(dims_a[m] for m in a_in), _prod(dims_a[m] for m in a_int))
    lat_b = (lat_a[1], _prod(dims_b[r] for r in b_out), lat_a[2])
    dims_c = [int(dims_a[m]) for m in a_out] + [int(dims_b[r]) for r in b_out] + [int(dims_a[m]) for m in a_int]
    labels_c = "".join([la[m] for m in a_out] + [lb[r] for r in b_out] + [la[m] for m in a_int])
    return seq_a, seq_b, lat_a, lat_b, dims_c, labels_c, (a_in, a_out, b_out)


class Contraction(ctypes.Structure):
    """``lnt_contraction`` of include/lnt_batch.h."""

    _fields_ = [
        ("order_a", ctypes.c_int32), ("order_b", ctypes.c_int32),
        ("dims_a", ctypes.c_int64 * MAX_ORDER), ("dims_b", ctypes.c_int64 * MAX_ORDER),
        ("cur_seq_a", ctypes.c_int32 * MAX_ORDER), ("new_seq_a", ctypes.c_int32 * MAX_ORDER),
        ("cur_seq_b", ctypes.c_int32 * MAX_ORDER), ("new_seq_b", ctypes.c_int32 * MAX_ORDER),
        ("lat_a", ctypes.c_int64 * 3), ("lat_b", ctypes.c_int64 * 3),
        ("nnz_a", ctypes.c_int64), ("nnz_b", ctypes.c_int64),
    ]


class Instance:
    """One contraction of a batch: ``C = A(labels_a) * B(labels_b)`` with A's
    keys linearised in ``cur_seq_a`` (default order when None), likewise B."""

    def __init__(self, dims_a, labels_a, dims_b, labels_b, nnz_a, nnz_b, cur_seq_a=None, cur_seq_b=None):
        self.dims_a = [int(d) for d in dims_a]
        self.dims_b = [int(d) for d in dims_b]
        self.labels_a, self.labels_b = labels_a, labels_b
        self.nnz_a, self.nnz_b = int(nnz_a), int(nnz_b)
        self.cur_seq_a = list(range(len(self.dims_a))) if cur_seq_a is None else list(cur_seq_a)
        self.cur_seq_b = list(range(len(self.dims_b))) if cur_seq_b is None else list(cur_seq_b)
        plan = lattice_plan(self.dims_a, labels_a, self.dims_b, labels_b)
        self.seq_a, self.seq_b, self.lat_a, self.lat_b, self.dims_c, self.labels_c, _ = plan

    @property
    def dimensionality_a(self):
        return _prod(self.dims_a)

    @property
    def dimensionality_b(self):
        return _prod(self.dims_b)

    def to_struct(self) -> Contraction:
        c = Contraction()
        c.order_a, c.order_b = len(self.dims_a), len(self.dims_b)
        for m, d in enumerate(self.dims_a):
            c.dims_a[m] = d
        for m, d in enumerate(self.dims_b):
            c.dims_b[m] = d
        for m in range(len(self.dims_a)):
            c.cur_seq_a[m] = self.cur_seq_a[m]
            c.new_seq_a[m] = self.seq_a[m]
        for m in range(len(self.dims_b)):
            c.cur_seq_b[m] = self.cur_seq_b[m]
            c.new_seq_b[m] = self.seq_b[m]
        for q in range(3):
            c.lat_a[q] = self.lat_a[q]
            c.lat_b[q] = self.lat_b[q]
        c.nnz_a, c.nnz_b = self.nnz_a, self.nnz_b
        return c


def packed_offsets(counts: Sequence[int]) -> np.ndarray:
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(np.asarray(counts, dtype=np.int64), out=off[1:])
    return off


class Batch:
    """Owns an ``lnt_batch``.  Operands are packed device tensors: instance
    i's A entries at ``a_offsets[i]:a_offsets[i+1]``, likewise B."""

    def __init__(self, ctx, instances: List[Instance]):
        self.ctx = ctx
        self.instances = list(instances)
        arr = (Contraction * max(len(self.instances), 1))()
        for i, inst in enumerate(self.instances):
            arr[i] = inst.to_struct()
        self._h = _vp()
        check(lib().lnt_batch_create(ctx.handle, arr, len(self.instances), ctypes.byref(self._h)))
        self.a_offsets = packed_offsets([x.nnz_a for x in self.instances])
        self.b_offsets = packed_offsets([x.nnz_b for x in self.instances])
        self._out = None  # output tensors of the recorded graph

    @property
    def count(self):
        return len(self.instances)

    def close(self):
        if self._h:
            lib().lnt_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_param(self, name: str, value: int):
        check(lib().lnt_batch_set_param(self._h, name.encode(), int(value)))

    def get_param(self, name: str) -> int:
        v = ctypes.c_int64()
        check(lib().lnt_batch_get_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_accumulators(self, acc):
        """Accumulator of each instance (``dispatch.ACC_*``): ACC_DENSE and
        ACC_MTT sum the instance's multi-entry output columns in a dense LDS
        accumulator over its lattice rows (height <= 8192), any other value
        keeps the batch's merge and sort paths.  Results do not change."""
        a = np.ascontiguousarray(acc, dtype=np.int32)
        if a.shape != (self.count,):
            raise ValueError("one accumulator per instance is required")
        check(lib().lnt_batch_set_accumulators(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))

    @staticmethod
    def _dev(x):
        p, mem, _ = _ptr_mem(x)
        if mem != MEM_DEVICE:
            raise ValueError("batch operands must be device tensors")
        return p

    def contract(self, ak, av, bk, bv):
        """Returns (keys, values, offsets): C of instance i is
        keys[offsets[i]:offsets[i+1]], in the instance's C lattice order."""
        import torch

        n = ctypes.c_int64()
        check(lib().lnt_batch_symbolic(self._h, self._dev(ak), self._dev(av), self._dev(bk), self._dev(bv),
                                       ctypes.byref(n)))
        keys = torch.empty(max(n.value, 1), dtype=torch.int64, device=ak.device)
        vals = torch.empty(max(n.value, 1), dtype=torch.float64, device=ak.device)
        off = np.zeros(self.count + 1, dtype=np.int64)
        check(lib().lnt_batch_numeric(self._h, keys.data_ptr(), vals.data_ptr(), n.value, off.ctypes.data_as(_i64p)))
        return keys[: n.value], vals[: n.value], off

    def nprod(self, ak, bk) -> np.ndarray:
        out = np.zeros(max(self.count, 1), dtype=np.int64)
        check(lib().lnt_batch_nprod(self._h, self._dev(ak), self._dev(bk), out.ctypes.data_as(_i64p)))
        return out[: self.count]

    def capture(self, ak, av, bk, bv, capacity=None):
        """Records the pipeline into a hipGraph writing into tensors owned by
        this object (sized exactly, or to ``capacity``).  Returns nnz(C)."""
        import torch

        n = ctypes.c_int64()
        if capacity is None:
            # An unrecorded run with no output room gives the exact size.
            check(lib().lnt_batch_capture(self._h, self._dev(ak), self._dev(av), self._dev(bk), self._dev(bv),
                                          None, None, 0, ctypes.byref(n)))
            capacity = n.value
        keys = torch.empty(max(capacity, 1), dtype=torch.int64, device=ak.device)
        vals = torch.empty(max(capacity, 1), dtype=torch.float64, device=ak.device)
        check(lib().lnt_batch_capture(self._h, self._dev(ak), self._dev(av), self._dev(bk), self._dev(bv),
                                      keys.data_ptr(), vals.data_ptr(), capacity, ctypes.byref(n)))
        self._out = (keys, vals, (ak, av, bk, bv))
        return n.value

    def replay(self):
        """Runs the recorded graph; returns (keys, values, offsets) views."""
        off = np.zeros(self.count + 1, dtype=np.int64)
        check(lib().lnt_batch_replay(self._h, off.ctypes.data_as(_i64p)))
        n = int(off[-1])
        return self._out[0][:n], self._out[1][:n], off

    def launch(self):
        check(lib().lnt_batch_launch(self._h))
