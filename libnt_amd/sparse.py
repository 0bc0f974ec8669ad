"""Host-side mirror of LibMIA's SparseMIA index-expression front end.

``SparseMIA`` mirrors ``SparseMIA<double,N>`` (multi_index_array/SparseMIA.h:
167-773): dims, int64 linear indices and double values (here GPU tensors), the
mode order of the indices (``linIdxSequence``, SparseMIABase.h:98-153) and a
sorted flag.  ``A("ijkl")`` mirrors ``A(i,j,k,l)`` (MIA.h:222-229) and returns
an ``Atom`` on which ``+``, ``-``, ``*`` and ``~`` follow the reference's
expression rules (MIA_Expr.h).  A label is a letter preceded by its
element-wise marks: ``A("i!jk!!l")`` mirrors ``A(i,!j,k,!!l)`` (Index.h:38-57);
an upper-case letter is one mark (``"iJ"`` = ``"i!j"``).  Every operation runs
on the GPU through the C-ABI (include/lnt.h): change_sort_order, merge,
lattice product.
"""
from __future__ import annotations

import numpy as np

from . import OP_ADD, OP_HADAMARD, OP_SUB, SPGEMM_AUTO, Context

_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _prod(xs):
    p = 1
    for x in xs:
        p *= int(x)
    return p


def parse_labels(labels, allow_repeats: bool = False):
    """``"i!jk!!l"`` -> ``[("i", 0), ("j", 1), ("k", 0), ("l", 2)]``: one
    (letter, element-wise count) per mode; an upper-case letter counts one
    mark.  A list of such pairs passes through.  A letter may repeat only
    where `allow_repeats` (indexing a tensor: unary contraction/attraction)."""
    if not isinstance(labels, str):
        return [(str(c), int(b)) for c, b in labels]
    out, bangs = [], 0
    for ch in labels:
        if ch == "!":
            bangs += 1
            continue
        if not ch.isalpha():
            raise ValueError("index labels are letters, optionally preceded by '!': %r" % labels)
        out.append((ch.lower(), bangs + (1 if ch.isupper() else 0)))
        bangs = 0
    if bangs:
        raise ValueError("a trailing '!' in %r" % labels)
    letters = [c for c, _ in out]
    if len(set(letters)) != len(letters) and not allow_repeats:
        raise ValueError("an index label repeats within one operand: %r" % labels)
    return out


def _plan_string(tokens) -> str:
    """Plain labels lower case, element-wise ones upper case (the form the
    lattice planner reads)."""
    return "".join(c.upper() if b else c for c, b in tokens)


class SparseMIA:
    """Sparse order-N tensor: keys (int64) and values (float64) on the GPU."""

    def __init__(self, dims, keys=None, vals=None, seq=None, is_sorted=False, ctx=None):
        import torch

        self.dims = tuple(int(d) for d in dims)
        self.ctx = ctx or default_context()
        if keys is None:
            keys = torch.empty(0, dtype=torch.int64, device="cuda")
            vals = torch.empty(0, dtype=torch.float64, device="cuda")
        if isinstance(keys, np.ndarray):
            keys = torch.from_numpy(np.ascontiguousarray(keys, dtype=np.int64)).cuda()
        if isinstance(vals, np.ndarray):
            vals = torch.from_numpy(np.ascontiguousarray(vals, dtype=np.float64)).cuda()
        if keys.shape != vals.shape:
            raise ValueError("keys and values differ in length")
        self.keys = keys.contiguous()
        self.vals = vals.contiguous()
        self.seq = list(range(len(self.dims))) if seq is None else [int(x) for x in seq]
        self.is_sorted = bool(is_sorted)

    # ------------------------------------------------------------ properties
    @property
    def order(self) -> int:
        return len(self.dims)

    @property
    def dimensionality(self) -> int:
        return _prod(self.dims)

    @property
    def nnz(self) -> int:
        return int(self.keys.shape[0])

    def linIdxSequence(self):
        return list(self.seq)

    # ----------------------------------------------------------------- sorts
    def sort(self, seq=None):
        """SparseMIABase::sort() / sort(seq) (SparseMIABase.h:590-645, 701-734)."""
        new = self.seq if seq is None else [int(x) for x in seq]
        if self.nnz:
            self.ctx.change_sort_order(self.keys, self.vals, self.dims, self.seq, new, self.is_sorted)
        self.seq = list(new)
        self.is_sorted = True
        return self

    def change_linIdx_sequence(self, seq):
        """Relinearise without sorting (SparseMIABase.h:768-814)."""
        seq = [int(x) for x in seq]
        if seq == self.seq or self.order == 1:
            return self
        if self.nnz:
            self.ctx.change_linidx(self.keys, self.dims, self.seq, seq)
        self.seq = seq
        self.is_sorted = False
        return self

    def copy(self):
        return SparseMIA(self.dims, self.keys.clone(), self.vals.clone(), self.seq, self.is_sorted, self.ctx)

    # ------------------------------------------------------------ comparison
    def _compare(self, other: "SparseMIA", precision):
        # SparseMIABase.h:1243-1296: dims must agree; this is sorted, the other
        # sorted into this one's sequence (both in place), then compared.
        if tuple(self.dims) != tuple(other.dims):
            return False
        self.sort()
        other.sort(self.seq)
        return self.ctx.compare(self.keys, self.vals, other.keys, other.vals, precision)

    def __eq__(self, other):
        """operator== (SparseMIABase.h:1616-1623): exact values; entries with
        |value| <= 1.11e-16 count as absent."""
        if not isinstance(other, SparseMIA):
            return NotImplemented
        return self._compare(other, None)

    __hash__ = None

    def fuzzy_equals(self, other: "SparseMIA", precision: float) -> bool:
        """fuzzy_equals (SparseMIABase.h:1627-1635): isEqualFuzzy per pair."""
        return self._compare(other, precision)

    # -------------------------------------------------------------- helpers
    def __call__(self, labels):
        """A("ijk") -> Atom.  A repeated letter contracts (plain) or attracts
        (``!``) at once (perform_unary, MIA_Expr.h:695-731); contracting every
        mode gives a float."""
        tokens = parse_labels(labels, allow_repeats=True)
        if len(tokens) != self.order:
            raise ValueError("expected %d index labels, got %r" % (self.order, labels))
        if len({c for c, _ in tokens}) == len(tokens):
            return Atom(self, tokens)
        return self._unary(tokens)

    def _unary(self, tokens):
        """SparseMIABase::contract_attract (SparseMIABase.h:1913-2053).  Groups
        are numbered as the reference's expression code pulls them:
        contractions, then attractions, each in the order of the last
        occurrence of its letter (ExprUtil.h:120-240), which fixes the order of
        the sums.  The result's labels: the other letters in mode order, then
        the attracted ones."""
        from . import UNARY_ATTRACT, UNARY_CONTRACT

        last = {c: p for p, (c, _) in enumerate(tokens)}
        marks = {}
        for c, b in tokens:
            if marks.setdefault(c, b) != b:
                raise ValueError("repeated index %r has different element-wise counts" % c)
        letters = [c for c, _ in tokens]
        rep = [c for c in sorted(last, key=last.get) if letters.count(c) > 1]
        con = [c for c in rep if marks[c] == 0]
        att = [c for c in rep if marks[c] > 0]
        number = {c: g + 1 for g, c in enumerate(con + att)}
        group = [number.get(c, 0) for c in letters]
        kind = [UNARY_CONTRACT] * len(con) + [UNARY_ATTRACT] * len(att)
        ck, cv = self.ctx.contract_attract(self.keys, self.vals, self.dims, self.seq, group, kind)
        stays = [m for m in range(self.order) if group[m] == 0]
        dims = [self.dims[m] for m in stays] + [self.dims[letters.index(c)] for c in att]
        out = [tokens[m] for m in stays] + [(c, marks[c]) for c in att]
        if not dims:
            return float(cv[0].item()) if cv.numel() else 0.0
        return Atom(SparseMIA(dims, ck, cv, None, True, self.ctx), out)

    def to_dense(self):
        """Dense numpy copy in the default (first mode fastest) layout."""
        import torch

        k = self.keys.clone()
        if self.seq != list(range(self.order)) and self.nnz:
            self.ctx.change_linidx(k, self.dims, self.seq, list(range(self.order)))
        d = np.zeros(self.dimensionality)
        d[k.cpu().numpy()] = self.vals.cpu().numpy()
        return d.reshape(self.dims, order="F")

    def default_order_arrays(self):
        """(keys, values) as numpy, sorted in the default mode order."""
        c = self.copy().sort(list(range(self.order)))
        return c.keys.cpu().numpy(), c.vals.cpu().numpy()


class DenseMIA:
    """Dense order-N tensor as the dense operand of sparse x dense products:
    dims and the elements, first mode fastest (the reference's DenseMIA
    layout, multi_index_array/DenseMIA.h), as a 1-D GPU tensor."""

    def __init__(self, dims, data, ctx=None):
        import torch

        self.dims = tuple(int(d) for d in dims)
        self.ctx = ctx or default_context()
        if isinstance(data, np.ndarray):
            if data.ndim != 1:
                raise ValueError("data is the flat element array, first mode fastest")
            data = torch.from_numpy(np.ascontiguousarray(data, dtype=np.float64)).cuda()
        if data.dim() != 1 or data.numel() != _prod(self.dims):
            raise ValueError("data must hold prod(dims) elements in one dimension")
        self.data = data.contiguous()

    @classmethod
    def from_array(cls, array: np.ndarray, ctx=None) -> "DenseMIA":
        """From an N-d numpy array indexed [i0, i1, ...]."""
        return cls(array.shape, np.ravel(array, order="F"), ctx)

    def __call__(self, labels) -> "DenseAtom":
        tokens = parse_labels(labels)
        if len(tokens) != len(self.dims):
            raise ValueError("expected %d index labels, got %r" % (len(self.dims), labels))
        return DenseAtom(self, tokens)


class DenseAtom:
    """B(i,j,...) for a DenseMIA B: the right operand of Atom * DenseAtom."""

    def __init__(self, dense: DenseMIA, tokens):
        self.mia = dense
        self.tokens = parse_labels(tokens)


class Atom:
    """A(i,j,...) -- a tensor with index labels (MIA_Atom, MIA_Expr.h:663-689).

    ``tokens`` holds (letter, element-wise count) per mode; ``labels`` is the
    same as a string, element-wise letters upper case; ``inter_count`` is the
    number of inter indices a product left at the back, the ones ``~``
    lowers (the reference's inter_product_number, MIA_Expr.h:903-920)."""

    def __init__(self, mia: SparseMIA, labels, inter_count: int = 0):
        self.mia = mia
        self.tokens = parse_labels(labels)
        self.labels = _plan_string(self.tokens)
        self.inter_count = int(inter_count)

    def __invert__(self) -> "Atom":
        """~(A*B): the product's inter indices lose one mark (MIA_Expr.h:
        257-301, 1376): ``!j`` becomes a plain index of the next product,
        ``!!l`` an inter index again.  No data moves."""
        t = list(self.tokens)
        for q in range(len(t) - self.inter_count, len(t)):
            t[q] = (t[q][0], max(t[q][1] - 1, 0))
        return Atom(self.mia, t, 0)

    # ------------------------------------------------------------- merges
    def _merge(self, other: "Atom", op: int) -> SparseMIA:
        """outside_merge (SparseMIABase.h:1357-1455): outside_scanMerge when
        either operand is sorted, sortMerge when neither is."""
        a, b = self.mia, other.mia
        if sorted(self.labels) != sorted(other.labels) or any(bg for _, bg in self.tokens + other.tokens):
            raise ValueError("merge operands must carry the same plain labels")
        if b is a:
            b = b.copy()  # SparseMIABase.h:1367-1371
        index_order = [other.labels.index(c) for c in self.labels]  # A mode -> B mode
        for m, r in enumerate(index_order):
            if a.dims[m] != b.dims[r]:
                raise ValueError("merge dimensions mismatch")
        b_in_a_order = [index_order[m] for m in a.seq]
        lhs_order = [self.labels.index(c) for c in other.labels]  # B mode -> A mode
        if b.is_sorted and a.is_sorted:
            if b.nnz < a.nnz:
                b.sort(b_in_a_order)
            else:
                a.sort([lhs_order[r] for r in b.seq])
        elif b.is_sorted:
            a.sort([lhs_order[r] for r in b.seq])
        elif a.is_sorted:
            b.sort(b_in_a_order)
        else:
            # Neither sorted: C = copy of A, C.sortMerge(B) (SparseMIABase.h:
            # 1393-1398, SparseMIA.h:833-887): the smaller operand is
            # relinearised into the other's sequence, then concatenation, sort
            # and collect_duplicates (keys whose entries are all zero vanish).
            if a.nnz > b.nnz:
                b.change_linIdx_sequence(b_in_a_order)
                ak, seq = a.keys, a.seq
            else:
                seq = [lhs_order[r] for r in b.seq]
                c = a.copy()
                c.change_linIdx_sequence(seq)
                ak = c.keys
            ck, cv = a.ctx.sort_merge(ak, a.vals, b.keys, b.vals, op, a.dimensionality)
            return SparseMIA(a.dims, ck.clone(), cv.clone(), seq, True, a.ctx)
        ck, cv = a.ctx.merge(a.keys, a.vals, b.keys, b.vals, op)
        return SparseMIA(a.dims, ck.clone(), cv.clone(), a.seq, True, a.ctx)

    def __add__(self, other):
        return self._merge(other, OP_ADD)

    def __sub__(self, other):
        return self._merge(other, OP_SUB)

    def _merge_into_self(self, other: "Atom", op: int) -> "Atom":
        """Destructive A(...) += / -= B(...) (SparseMIA::merge,
        SparseMIA.h:1030-1050): the merge, written back into this tensor."""
        c = self._merge(other, op)
        m = self.mia
        m.keys, m.vals, m.seq, m.is_sorted = c.keys, c.vals, list(c.seq), c.is_sorted
        return self

    def __iadd__(self, other):
        return self._merge_into_self(other, OP_ADD)

    def __isub__(self, other):
        return self._merge_into_self(other, OP_SUB)

    # ----------------------------------------------------------- products
    def plan(self, other: "Atom"):
        """Lattice bookkeeping of MIA_Expr.h:449-561 (see ``batch.lattice_plan``):
        A rows = outer, columns = inner (contracted), tabs = inter, each in A's
        mode order; B rows = inner matched to A's order, columns = B's outer
        modes, tabs = inter matched.  C's modes are [A outer, B outer, inter]."""
        from .batch import lattice_plan

        return lattice_plan(self.mia.dims, self.labels, other.mia.dims, other.labels)

    def _mul_dense(self, other: DenseAtom) -> "Atom":
        """Sparse x dense with no contracted index (SparseMIABase::noLatticeMult
        with a DenseMIA, SparseMIABase.h:1829-1910): A sorted to [outer, inter]
        order, then C(l, r, t) = A(l, t) * B(r, t) for non-zero factors.  C's
        modes are [A outer, B outer, inter] and its keys run B outer fastest,
        then A outer, then inter (:1899)."""
        a, dense = self.mia, other.mia
        ta, tb = self.tokens, other.tokens
        la, lb = [c for c, _ in ta], [c for c, _ in tb]
        for c, bang in ta:
            if c in lb:
                if bang == 0 or dict(tb)[c] != bang:
                    raise ValueError("index %r: a sparse x dense product takes only element-wise shared indices" % c)
            elif bang:
                raise ValueError("inter index %r appears in one operand only" % c)
        if any(bang and c not in la for c, bang in tb):
            raise ValueError("inter index appears in one operand only")
        a_out = [m for m, c in enumerate(la) if c not in lb]
        a_int = [m for m, c in enumerate(la) if c in lb]
        b_out = [r for r, c in enumerate(lb) if c not in la]
        b_int = [lb.index(la[m]) for m in a_int]
        for m, r in zip(a_int, b_int):
            if a.dims[m] != dense.dims[r]:
                raise ValueError("matched indices differ in size")
        a.sort(a_out + a_int)
        ck, cv = a.ctx.sparse_dense_mult(a.keys, a.vals, _prod(a.dims[m] for m in a_out), dense.data, dense.dims,
                                         b_out, b_int)
        nl, nr, ni = len(a_out), len(b_out), len(a_int)
        dims = [a.dims[m] for m in a_out] + [dense.dims[r] for r in b_out] + [a.dims[m] for m in a_int]
        seq = list(range(nl, nl + nr)) + list(range(nl)) + list(range(nl + nr, nl + nr + ni))
        tokens = [ta[m] for m in a_out] + [tb[r] for r in b_out] + [ta[m] for m in a_int]
        return Atom(SparseMIA(dims, ck, cv, seq, True, a.ctx), tokens, ni)

    def __mul__(self, other: "Atom"):
        """The product; a complete inner product (every index contracted)
        returns its value as a float (sparse_mia_mult_test.cpp:159-165)."""
        if isinstance(other, DenseAtom):
            return self._mul_dense(other)
        a, b = self.mia, other.mia
        if b is a:
            b = b.copy()
        counts_a, counts_b = dict(self.tokens), dict(other.tokens)
        for c in counts_a.keys() & counts_b.keys():
            # product_rule_different_elem_wise (Index.h:95-101)
            if counts_a[c] != counts_b[c]:
                raise ValueError("index %r has different element-wise counts on the two sides" % c)
        seq_a, seq_b, lat_a, lat_b, dims_c, labels_c, (a_in, a_out, b_out) = self.plan(Atom(b, other.tokens))
        a.sort(seq_a)
        b.sort(seq_b)
        if not a_in and not a_out and not b_out:
            # pure Hadamard (SparseMIABase.h:1717-1826 with no outer indices)
            ck, cv = a.ctx.merge(a.keys, a.vals, b.keys, b.vals, OP_HADAMARD)
        else:
            ck, cv = a.ctx.lattice_mult(a.keys, a.vals, lat_a, b.keys, b.vals, lat_b, SPGEMM_AUTO)
        # toStraightLattice/toLatticeSort leave the operands flagged unsorted
        # (SparseMIABase.h:1507, SparseMIA.h:1063-1122).
        a.is_sorted = False
        b.is_sorted = False
        if not dims_c:
            return float(cv[0].item()) if cv.numel() else 0.0
        tokens_c = [(ch.lower(), counts_a.get(ch.lower(), counts_b.get(ch.lower()))) for ch in labels_c]
        c = SparseMIA(dims_c, ck.clone(), cv.clone(), None, True, a.ctx)
        return Atom(c, tokens_c, len(dims_c) - len(a_out) - len(b_out))

    # ---------------------------------------------------------- assignment
    def assign(self, labels) -> SparseMIA:
        """C(labels) = expr: relabels C's mode order without moving data
        (SparseMIA::assign, SparseMIA.h:891-928); indices match by letter,
        whatever their element-wise counts (MIA_Expr.h:112-135)."""
        mine = [c for c, _ in self.tokens]
        theirs = [c for c, _ in parse_labels(labels)]
        if sorted(theirs) != sorted(mine):
            raise ValueError("assignment labels differ from the expression's")
        perm = [mine.index(c) for c in theirs]  # new mode m <- old mode perm[m]
        old = self.mia
        dims = [old.dims[p] for p in perm]
        inv = {p: m for m, p in enumerate(perm)}
        seq = [inv[s] for s in old.seq]
        return SparseMIA(dims, old.keys, old.vals, seq, old.is_sorted, old.ctx)
