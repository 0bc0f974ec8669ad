"""Autotuner of the batched-contraction pipeline (``include/lnt_batch.h``).

Coordinate descent over the pipeline's launch parameters -- radix digit
width and tile shape of the sorts (context parameters ``digit_bits``,
``sort_block``, ``sort_items``), the small-segment bin threshold and work-item size
(``small_max``, ``item_products``), the row-window target of large segments
(``window``) and the persistent grid (``blocks_per_cu``) -- each candidate
timed on a fixed subsample of the batch.  Results never depend on the
parameters; only the time does.  No counterpart in the reference.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional

import numpy as np

from .batch import Batch, packed_offsets

SPACE = {
    "digit_bits": [8, 7, 6],
    "sort_block": [1024, 512, 256],
    "sort_items": [12, 16, 8],
    "small_max": [1024, 512, 256],
    "item_products": [1024, 512, 256],
    "window": [1280, 1024, 768, 1536],
    "blocks_per_cu": [2, 3, 4, 1],
}
CONTEXT_PARAMS = ("digit_bits", "sort_block", "sort_items")


def subsample(instances, ops, a_off, b_off, stride: int):
    """Every `stride`-th instance with its packed operands."""
    import torch

    idx = list(range(0, len(instances), stride))
    ak, av, bk, bv = ops
    sub = [instances[i] for i in idx]
    cat = lambda t, off: torch.cat([t[off[i]:off[i + 1]] for i in idx]) if idx else t[:0]
    return sub, (cat(ak, a_off), cat(av, a_off), cat(bk, b_off), cat(bv, b_off))


def _apply(ctx, batch: Batch, cfg: Dict[str, int]):
    for k, v in cfg.items():
        if k in CONTEXT_PARAMS:
            ctx.set_param(k, v)
        else:
            batch.set_param(k, v)


def _time(batch: Batch, ops, reps: int) -> float:
    import torch

    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ck, cv, off = batch.contract(*ops)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
        del ck, cv
    return best * 1e3


def tune(ctx, instances, ops, stride: int = 5, passes: int = 2, reps: int = 2, log=None) -> Dict[str, object]:
    """Tunes on every `stride`-th instance; returns {"config", "ms", "default_ms",
    "evaluations"}.  The context's sort parameters are left at the best values;
    apply the batch parameters with ``apply_batch``."""
    a_off = packed_offsets([x.nnz_a for x in instances])
    b_off = packed_offsets([x.nnz_b for x in instances])
    sub, sub_ops = subsample(instances, ops, a_off, b_off, stride)
    b = Batch(ctx, sub)
    cfg = {k: v[0] for k, v in SPACE.items()}
    _apply(ctx, b, cfg)
    _time(b, sub_ops, 1)  # warm-up: workspaces
    best = _time(b, sub_ops, reps)
    default_ms = best
    evals = 1
    for p in range(passes):
        improved = False
        for name, values in SPACE.items():
            for v in values:
                if v == cfg[name]:
                    continue
                trial = dict(cfg, **{name: v})
                _apply(ctx, b, trial)
                t = _time(b, sub_ops, reps)
                evals += 1
                if log:
                    log("tune %-14s %5d: %.2f ms (best %.2f)" % (name, v, t, best))
                if t < best * 0.99:
                    best, cfg, improved = t, trial, True
            _apply(ctx, b, cfg)
        if not improved:
            break
    _apply(ctx, b, cfg)
    b.close()
    return {"config": cfg, "ms": best, "default_ms": default_ms, "evaluations": evals}


def apply_batch(ctx, batch: Batch, cfg: Dict[str, int]):
    _apply(ctx, batch, cfg)
