"""Synthetic workloads with fixed seeds.

``batched_contractions`` draws the batched-contraction mix: every instance is
an independent ``C = A * B`` of two sparse FP64 tensors of order 3 or 4 that
contracts 1 or 2 modes; dims are log-uniform integers in [8, 128], nnz per
operand log-uniform in [100, 50k] capped at half the operand's
dimensionality, coordinates uniform without replacement, values uniform in
[-1, 1).  The description of instance i comes from numpy's PCG64 seeded with
(seed, i) (portable), its keys and values from the counter-based device
generators (``lnt_testgen_batch``), so any instance can be rebuilt alone.
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from .batch import Instance, packed_offsets

CONTRACTED = "ab"
OUTER_A = "cdef"
OUTER_B = "ghij"

# Seeds of the value and key streams of the batched mix, per operand.
KEY_SEED_A, VAL_SEED_A = 0x4B41, 0x5641
KEY_SEED_B, VAL_SEED_B = 0x4B42, 0x5642


def _logu_int(rng, lo: int, hi: int, size=None):
    """Integers log-uniform in [lo, hi]."""
    x = np.exp(rng.uniform(math.log(lo), math.log(hi + 1), size))
    return np.minimum(np.floor(x).astype(np.int64), hi)


def batched_contractions(count: int = 10000, seed: int = 4, nnz_lo: int = 100, nnz_hi: int = 50000,
                         dim_lo: int = 8, dim_hi: int = 128) -> List[Instance]:
    out = []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        oa, ob = (int(x) for x in rng.integers(3, 5, size=2))
        c = int(rng.integers(1, 3))
        dc = [int(x) for x in _logu_int(rng, dim_lo, dim_hi, c)]
        da = [int(x) for x in _logu_int(rng, dim_lo, dim_hi, oa - c)]
        db = [int(x) for x in _logu_int(rng, dim_lo, dim_hi, ob - c)]
        labels_a = list(CONTRACTED[:c]) + list(OUTER_A[: oa - c])
        dims_a = dc + da
        labels_b = list(CONTRACTED[:c]) + list(OUTER_B[: ob - c])
        dims_b = dc + db
        pa = rng.permutation(oa)
        pb = rng.permutation(ob)
        labels_a = "".join(labels_a[p] for p in pa)
        dims_a = [dims_a[p] for p in pa]
        labels_b = "".join(labels_b[p] for p in pb)
        dims_b = [dims_b[p] for p in pb]
        cap_a = math.prod(dims_a) // 2
        cap_b = math.prod(dims_b) // 2
        na = min(int(_logu_int(rng, nnz_lo, nnz_hi)), cap_a)
        nb = min(int(_logu_int(rng, nnz_lo, nnz_hi)), cap_b)
        out.append(Instance(dims_a, labels_a, dims_b, labels_b, na, nb))
    return out


def generate_operands(ctx, instances: List[Instance], device="cuda"):
    """Packed device operands (ak, av, bk, bv) of a batch, default order."""
    import torch

    a_off = packed_offsets([x.nnz_a for x in instances])
    b_off = packed_offsets([x.nnz_b for x in instances])
    ak = torch.empty(max(int(a_off[-1]), 1), dtype=torch.int64, device=device)
    av = torch.empty(max(int(a_off[-1]), 1), dtype=torch.float64, device=device)
    bk = torch.empty(max(int(b_off[-1]), 1), dtype=torch.int64, device=device)
    bv = torch.empty(max(int(b_off[-1]), 1), dtype=torch.float64, device=device)
    ctx.testgen_batch(ak, av, a_off, [x.dimensionality_a for x in instances], KEY_SEED_A, VAL_SEED_A)
    ctx.testgen_batch(bk, bv, b_off, [x.dimensionality_b for x in instances], KEY_SEED_B, VAL_SEED_B)
    return ak[: int(a_off[-1])], av[: int(a_off[-1])], bk[: int(b_off[-1])], bv[: int(b_off[-1])]
