#!/usr/bin/env python3
"""Benchmark of the sparse multi-index-array hot path on one MI355X.

Default workload (the headline configuration, "sort/merge at scale"): two
order-4 sparse FP64 tensors over dims 256^4 (2^32 linear-index space), 50M
nonzeros each, supports overlapping 50% (75M distinct keys: 25M shared, 25M
A-only, 25M B-only), values U[-1,1), coordinates uniform.  One step =
change_sort_order of A and of B to a random mode permutation (uint64 keys +
FP64 payload), then A+B and A∘B.  Steps alternate between the permutation and
its way back so every step does the same work on sorted inputs.

Prints ONE JSON line.  value = sorted keys/s of the whole step
(nnz(A)+nnz(B) per step time).  roofline: the dominant kernel (the onesweep
radix pass) -- algorithmic bytes (32 B per key per pass) over its measured
average launch duration (HIP events on the library's stream during the timed
steps).  cpu_baseline: the CPU restatement of the reference algorithms
(oracle/), one core, on a bounded 1/8-scale sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
DIMS = [256, 256, 256, 256]
IDENT = [0, 1, 2, 3]
N_SHARED = 25_000_000
N_ONLY = 25_000_000
SEED_KEYS, SEED_A, SEED_B = 2024, 11, 12


def _perm(seed=7):
    return [int(x) for x in np.random.default_rng(seed).permutation(4)]


def make_c2(ctx, torch, scale=1):
    """A, B as GPU tensors, sorted in the default mode order."""
    ns, no = N_SHARED // scale, N_ONLY // scale
    na = nb = ns + no
    ak = torch.empty(na, dtype=torch.int64, device="cuda")
    bk = torch.empty(nb, dtype=torch.int64, device="cuda")
    av = torch.empty(na, dtype=torch.float64, device="cuda")
    bv = torch.empty(nb, dtype=torch.float64, device="cuda")
    # keys: one bijection of [0, 2^32); A = outputs [0, ns+no), B = [0, ns) + [ns+no, ns+2no)
    ctx.testgen_distinct_keys(ak, 1 << 32, SEED_KEYS, offset=0)
    ctx.testgen_distinct_keys(bk[:ns], 1 << 32, SEED_KEYS, offset=0)
    ctx.testgen_distinct_keys(bk[ns:], 1 << 32, SEED_KEYS, offset=ns + no)
    ctx.testgen_uniform(av, SEED_A)
    ctx.testgen_uniform(bv, SEED_B)
    ctx.sort_pairs(ak, av, 1 << 32)
    ctx.sort_pairs(bk, bv, 1 << 32)
    torch.cuda.synchronize()
    return ak, av, bk, bv


def cpu_baseline(scale=8, repeats=3):
    """The oracle (CPU restatement of the reference's algorithms) on one core:
    the same step on a 1/scale sample.  Returns the JSON object."""
    import oracle

    cpus = sorted(os.sched_getaffinity(0))
    old = set(cpus)
    os.sched_setaffinity(0, {cpus[0]})
    try:
        ns, no = N_SHARED // scale, N_ONLY // scale
        keys = oracle.gen_distinct_keys(ns + 2 * no, 1 << 32, SEED_KEYS)
        ak0 = keys[: ns + no].copy()
        bk0 = np.concatenate([keys[:ns], keys[ns + no:]])
        av0 = oracle.gen_uniform_vals(len(ak0), SEED_A)
        bv0 = oracle.gen_uniform_vals(len(bk0), SEED_B)
        ak0, av0 = oracle.radix_sort_pairs(ak0, av0, 1 << 32)
        bk0, bv0 = oracle.radix_sort_pairs(bk0, bv0, 1 << 32)
        perm = _perm()
        times = []
        for _ in range(repeats):
            ak, av, bk, bv = ak0.copy(), av0.copy(), bk0.copy(), bv0.copy()
            t0 = time.perf_counter()
            oracle.sort_seq_inplace(ak, av, DIMS, IDENT, perm, True)
            oracle.sort_seq_inplace(bk, bv, DIMS, IDENT, perm, True)
            oracle.merge(ak, av, bk, bv, oracle.OP_ADD)
            oracle.merge(ak, av, bk, bv, oracle.OP_HADAMARD)
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        return {
            "value": (len(ak0) + len(bk0)) / t,
            "unit": "keys/s",
            "cores": 1,
            "kind": "port",
            "sample": "1/%d-scale c2 step (%d+%d keys): 2x change_sort_order (relinearise + in-place MSD radix "
                      "2048/11-bit, thresholds 3000/50) + scan merge A+B + Hadamard; median of %d, %.2f s each"
                      % (scale, len(ak0), len(bk0), repeats, t),
        }
    finally:
        os.sched_setaffinity(0, old)


def _profiled_traffic():
    """HBM bytes per onesweep launch measured by rocprofv3 (FETCH_SIZE +
    WRITE_SIZE, scaled by the calibration kernels), from the committed profile
    summary; None if absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "onesweep_sort_merge_at_scale_pmc.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-scale", type=int, default=8)
    ap.add_argument("--calibrate", action="store_true",
                    help="after the timed steps, run a plain 1.6 GB write (fill) and read (sum) so that "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE can be scaled by kernels of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import torch

    import libnt_amd

    stream = torch.cuda.current_stream()
    ctx = libnt_amd.Context(0, stream=ctypes.c_void_p(stream.cuda_stream))
    ak, av, bk, bv = make_c2(ctx, torch)
    na, nb = ak.shape[0], bk.shape[0]
    cu_k = torch.empty(na + nb, dtype=torch.int64, device="cuda")
    cu_v = torch.empty(na + nb, dtype=torch.float64, device="cuda")
    ch_k = torch.empty(min(na, nb), dtype=torch.int64, device="cuda")
    ch_v = torch.empty(min(na, nb), dtype=torch.float64, device="cuda")
    perm = _perm()
    orders = [IDENT, perm]
    counts = {}

    def step(s):
        cur, new = orders[s % 2], orders[(s + 1) % 2]
        ctx.change_sort_order(ak, av, DIMS, cur, new, True)
        ctx.change_sort_order(bk, bv, DIMS, cur, new, True)
        counts["union"] = ctx.merge_into(ak, av, bk, bv, libnt_amd.OP_ADD, cu_k, cu_v)
        counts["hadamard"] = ctx.merge_into(ak, av, bk, bv, libnt_amd.OP_HADAMARD, ch_k, ch_v)

    s = 0
    for _ in range(args.warmup):
        step(s)
        s += 1
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(s)
        s += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = {k: ctx.profile(k) for k in ("histogram", "onesweep", "partition", "merge")}
    ctx.set_profiling(False)
    assert counts["union"] == N_SHARED + 2 * N_ONLY and counts["hadamard"] == N_SHARED, counts

    ms_per_step = elapsed * 1e3 / args.steps
    value = (na + nb) / (elapsed / args.steps)
    launches, ms, byts = prof["onesweep"]
    avg_ms = ms / max(launches, 1)
    achieved = byts / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    phases = {}
    for k, (n, t, b) in prof.items():
        phases[k] = {"launches": n, "avg_ms": round(t / max(n, 1), 4),
                     "GBps": round(b / (t * 1e-3) / 1e9, 1) if t > 0 and b > 0 else None}

    out = {
        "metric": "sorted keys/s (sort/merge at scale: change_sort_order of A and B + A+B + A∘B)",
        "value": value,
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64 keys + f64 values",
        "data": "synthetic (counter-based generator, fixed seeds)",
        "config": {
            "workload": "sort/merge at scale: order-4 dims 256^4, 50M nnz each, 50%% overlap, "
                        "change_sort_order to perm %s, then A+B (75M out) and A∘B (25M out)" % (perm,),
            "nnz_a": na, "nnz_b": nb, "union_nnz": counts["union"], "hadamard_nnz": counts["hadamard"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "onesweep radix pass (8-bit digit, 256x16 tile)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "avg_launch_ms": round(avg_ms, 4),
            "algorithmic_bytes_per_launch": byts / max(launches, 1),
            "traffic": None,
        },
        "phases": phases,
    }
    traffic = _profiled_traffic()
    if traffic:
        out["roofline"]["traffic"] = traffic
    if args.calibrate:
        cal = torch.empty(200_000_000, dtype=torch.float64, device="cuda")
        cal.fill_(1.0)
        float(cal.sum())
        torch.cuda.synchronize()
        del cal
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scale=args.cpu_scale)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
