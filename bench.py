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
(oracle/; RadixShuffle for the sorted inputs, as the reference sorts them),
one pinned core, one run of the same step on host copies of the same inputs.

Further objects on the same line (skip with --no-extras; --section runs one):
  small_contraction: C(i,j,m,n) = A(i,j,k,l)*B(k,l,m,n), dims 32^4 at 1%
    (10,486 nnz each), every strategy, plus change_sort_order of both
    operands to a seeded mode permutation; CPU baseline beside it;
  contraction: the large skewed contraction A(i,j,k)*B(k,l,m), A 4096x4096x512
    (858,993 nnz), B 512x4096x64 (13,422 nnz), i and l Zipf(1.1): whole
    contraction (unfold sorts + lattice product + refold) per accumulator,
    flops/s = 2 * scalar products / time, roofline over the compulsory bytes
    16 * (nnz(A) + nnz(B) + nnz(C)), CPU baseline on the same inputs;
  batched: 10,000 independent contractions (order 3-4, 1-2 contracted modes,
    dims log-U[8,128], nnz log-U[100,50k]) as one recorded hipGraph with
    autotuned parameters, against one lnt_contract_into call per instance
    (output into one reused buffer) with the restated polyalgorithm's
    strategy, the fixed default strategy and the learned model's strategy;
    roofline over the compulsory bytes; CPU baseline on a sample of instances;
  learned: the shipped cost model on the committed held-out timing set.
With --host-buffers, also "host_buffers": the same step with every call handed
host arrays (pinned, then pageable), staged through HBM by the library, so
the rate includes the PCIe copies (never the headline value).
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


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class _OneCore:
    """Pins this process to its first allowed core while alive (the
    reference's sparse path is single-threaded)."""

    def __enter__(self):
        self.old = set(os.sched_getaffinity(0))
        os.sched_setaffinity(0, {min(self.old)})
        return len(self.old)

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.old)


def _timed(fn, repeats):
    """Median wall time of `repeats` calls of fn (seconds)."""
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def cpu_baseline(ak0, av0, bk0, bv0, scale=1):
    """The oracle (CPU restatement of the reference's algorithms) on one
    pinned core: one run of the benchmark step on host arrays.  scale 1: host
    copies of the GPU's own (sorted) inputs; scale s > 1: the generator's
    first 1/s of the workload instead.  Returns the JSON object."""
    import oracle

    if scale > 1:
        ns, no = N_SHARED // scale, N_ONLY // scale
        keys = oracle.gen_distinct_keys(ns + 2 * no, 1 << 32, SEED_KEYS)
        ak0, av0 = oracle.radix_sort_pairs(keys[: ns + no], oracle.gen_uniform_vals(ns + no, SEED_A), 1 << 32)
        bk0, bv0 = oracle.radix_sort_pairs(np.concatenate([keys[:ns], keys[ns + no:]]),
                                           oracle.gen_uniform_vals(ns + no, SEED_B), 1 << 32)
        del keys
    perm = _perm()
    with _OneCore() as nproc:
        ak, av, bk, bv = ak0.copy(), av0.copy(), bk0.copy(), bv0.copy()
        t0 = time.perf_counter()
        oracle.sort_seq_shuffle_inplace(ak, av, DIMS, IDENT, perm, True)
        oracle.sort_seq_shuffle_inplace(bk, bv, DIMS, IDENT, perm, True)
        nu = len(oracle.merge(ak, av, bk, bv, oracle.OP_ADD)[0])
        nh = len(oracle.merge(ak, av, bk, bv, oracle.OP_HADAMARD)[0])
        t = time.perf_counter() - t0
    assert nu == len(ak0) + len(bk0) - nh
    assert bool(np.all(np.diff(ak) > 0)) and bool(np.all(np.diff(bk) > 0))
    what = ("the same inputs as the GPU step (host copies), full size" if scale == 1 else
            "a 1/%d-scale sample from the same generators" % scale)
    return {
        "value": (len(ak0) + len(bk0)) / t,
        "unit": "keys/s",
        "cores": 1,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "nproc": nproc,
        "sample": "one step on %s (%d+%d keys), one pinned core, single run of %.2f s: 2x change_sort_order "
                  "of sorted input by the reference's path (relinearise + RadixShuffle::permute, stages from "
                  "setupPermute) + scan merge A+B + Hadamard (the reference's reserve(nA*nB/2^Inter) replaced by "
                  "min(nA, nB))"
                  % (what, len(ak0), len(bk0), t),
    }


def _profiled_traffic(name="onesweep_sort_merge_at_scale_pmc.md", key="traffic_bytes_per_launch"):
    """HBM bytes measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by the
    calibration kernels), from a committed profile summary (per onesweep
    launch by default); None if absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def _log(msg):
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def bench_unsorted_sort(ctx, torch, n=50_000_000, reps=3):
    """A full radix sort of n distinct keys in generator (unsorted) order over
    2^32, values following: every digit pass runs (4 at 8 bits).  Median of
    `reps` sorts, each of freshly generated input (generation untimed)."""
    k = torch.empty(n, dtype=torch.int64, device="cuda")
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    times, passes = [], 0
    for r in range(reps + 1):
        ctx.testgen_distinct_keys(k, 1 << 32, 31 + r)
        ctx.testgen_uniform(v, 41 + r)
        torch.cuda.synchronize()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        ctx.sort_pairs(k, v, 1 << 32)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        passes = ctx.profile("onesweep")[0]
        ctx.set_profiling(False)
        if r:
            times.append(el)
    ok = bool((k[1:] > k[:-1]).all())
    t = sorted(times)[len(times) // 2]
    del k, v
    torch.cuda.empty_cache()
    return {"keys": n, "radix_passes": passes, "ms": round(t * 1e3, 3), "keys_per_s": n / t, "sorted": ok}


def _nprod_of(ak, bk, a_cols_div, b_rows_mod):
    """Exact scalar-product count sum_k nnzA(:,k) * nnzB(k,:) of a lattice
    pair whose A column is key // a_cols_div and B row key % b_rows_mod."""
    ca = np.bincount(ak // a_cols_div, minlength=b_rows_mod).astype(np.int64)
    cb = np.bincount(bk % b_rows_mod, minlength=b_rows_mod).astype(np.int64)
    return int((ca[:b_rows_mod] * cb).sum())


def _roofline(bytes_, seconds, what):
    gbs = bytes_ / seconds / 1e9
    return {"bound": "hbm", "kernel": what, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": bytes_, "traffic": None}


def _contract_once(ctx, A, la, B, lb, strategy):
    """nnz(C) of one contraction (device result freed)."""
    from libnt_amd import dispatch

    C, _, _ = dispatch.contract(ctx, A, la, B, lb, strategy)
    return C.nnz


def bench_small_contraction(ctx, torch, reps=20, cpu_repeats=5):
    """C(i,j,m,n) = A(i,j,k,l)*B(k,l,m,n) over dims 32^4 at 1% (10,486 nnz each,
    keys from the oracle's generators, unsorted), every strategy; and
    change_sort_order of both operands to a seeded permutation.  The CPU
    baseline runs the oracle's contraction and sort(seq) on the same arrays."""
    import oracle
    from libnt_amd import dispatch
    from libnt_amd.sparse import SparseMIA

    dims, n = [32, 32, 32, 32], 10486
    ak, av = oracle.gen_distinct_keys(n, 1 << 20, 1), oracle.gen_uniform_vals(n, 1001)
    bk, bv = oracle.gen_distinct_keys(n, 1 << 20, 2), oracle.gen_uniform_vals(n, 1002)
    A = SparseMIA(dims, torch.from_numpy(ak).cuda(), torch.from_numpy(av).cuda(), ctx=ctx)
    B = SparseMIA(dims, torch.from_numpy(bk).cuda(), torch.from_numpy(bv).cuda(), ctx=ctx)
    nprod = _nprod_of(ak, bk, 1024, 1024)  # A's (k,l) = key >> 10, B's (k,l) = key & 1023
    ms = dispatch.time_strategies(ctx, A, "ijkl", B, "klmn", (1 << dispatch.NUM_STRATEGIES) - 1, reps)
    best = int(np.argmin(np.where(ms > 0, ms, np.inf)))
    nnz_c = _contract_once(ctx, A, "ijkl", B, "klmn", best)
    perm = _perm(1)
    t_sort = []
    for _ in range(reps + 1):
        k1, v1 = A.keys.clone(), A.vals.clone()
        k2, v2 = B.keys.clone(), B.vals.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.change_sort_order(k1, v1, dims, IDENT, perm, False)
        ctx.change_sort_order(k2, v2, dims, IDENT, perm, False)
        torch.cuda.synchronize()
        t_sort.append(time.perf_counter() - t0)
    t_sort = float(np.median(t_sort[1:]))
    t_gpu = ms[best] * 1e-3
    res = {"workload": "C(i,j,m,n)=A(i,j,k,l)*B(k,l,m,n), dims 32^4, 10486 nnz each (1%%), unsorted; "
                       "change_sort_order of both to perm %s" % (perm,),
           "products": nprod, "nnz_c": nnz_c, "unit": "GFLOP/s (2 flops per scalar product)",
           "strategy_ms": {dispatch.strategy_name(s): round(float(ms[s]), 4) for s in range(len(ms)) if ms[s] > 0},
           "best": dispatch.strategy_name(best), "best_ms": round(float(ms[best]), 4),
           "gflops": round(2 * nprod / t_gpu / 1e9, 3),
           "sort_ms": round(t_sort * 1e3, 4), "sorted_keys_per_s": 2 * n / t_sort,
           "roofline": _roofline(16.0 * (2 * n + nnz_c), t_gpu, "whole contraction (%s), compulsory bytes"
                                 % dispatch.strategy_name(best))}
    with _OneCore():
        t_cc = _timed(lambda: oracle.contract(ak, av, dims, IDENT, False, "ijkl", bk, bv, dims, IDENT, False,
                                              "klmn"), cpu_repeats)
        t_cs = _timed(lambda: (oracle.sort_seq(ak, av, dims, IDENT, perm, False),
                               oracle.sort_seq(bk, bv, dims, IDENT, perm, False)), cpu_repeats)
    res["cpu_baseline"] = {"value": 2 * nprod / t_cc / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "port",
                           "ms": round(t_cc * 1e3, 3), "sort_ms": round(t_cs * 1e3, 3),
                           "sorted_keys_per_s": 2 * n / t_cs,
                           "sample": "the same arrays, one pinned core, median of %d: csc_no_accum contraction "
                                     "(unfold sorts, per-column expand/radix sort/collect) and 2x sort(seq)"
                                     % cpu_repeats}
    _log("small contraction: %s" % res)
    return res


def bench_contraction(ctx, torch, cpu_repeats=3, mask=None):
    """Large skewed contraction, every accumulator, timed whole; roofline over
    the compulsory bytes; the oracle on one core on the same arrays."""
    import oracle
    from libnt_amd import dispatch

    sp = dispatch.Spec("skewed", [4096, 4096, 512], "ijk", [512, 4096, 64], "klm", 858993, 13422,
                       "default", "default", (0,), (1,))
    A, B = dispatch.materialize(ctx, sp, 31)
    ka, va = A.keys.cpu().numpy(), A.vals.cpu().numpy()
    kb, vb = B.keys.cpu().numpy(), B.vals.cpu().numpy()
    nprod = _nprod_of(ka, kb, 4096 * 4096, 512)  # A's k = key // (4096*4096), B's k = key % 512
    ms = dispatch.time_strategies(ctx, A, sp.la, B, sp.lb, mask or (1 << dispatch.NUM_STRATEGIES) - 1, 3)
    timed = [s for s in range(len(ms)) if ms[s] > 0]
    best = min(timed, key=lambda s: ms[s])
    nnz_c = _contract_once(ctx, A, sp.la, B, sp.lb, best)
    res = {"workload": "A(i,j,k)*B(k,l,m), A 4096x4096x512 nnz 858993, B 512x4096x64 nnz 13422, "
                       "i and l Zipf(1.1); unfold sorts + lattice product + refold",
           "products": nprod, "nnz_c": nnz_c, "unit": "GFLOP/s (2 flops per scalar product)"}
    for s in timed:
        res[dispatch.strategy_name(s)] = {"ms": round(float(ms[s]), 3),
                                          "gflops": round(2 * nprod / (ms[s] * 1e-3) / 1e9, 2)}
    res["best"] = dispatch.strategy_name(best)
    res["gflops"] = round(2 * nprod / (ms[best] * 1e-3) / 1e9, 2)
    res["roofline"] = _roofline(16.0 * (A.nnz + B.nnz + nnz_c), ms[best] * 1e-3,
                                "whole contraction (%s), compulsory bytes 16*(nnzA+nnzB+nnzC)"
                                % dispatch.strategy_name(best))
    # HBM bytes of one contraction, measured by rocprofv3 on this workload.
    res["roofline"]["traffic"] = _profiled_traffic("traffic_contraction_skewed_4096x4096x512.md",
                                                   "traffic_bytes_per_contraction")
    if cpu_repeats < 1:
        _log("contraction: %s" % res)
        return res
    with _OneCore():
        t = _timed(lambda: oracle.contract(ka, va, sp.dims_a, IDENT[:3], True, sp.la, kb, vb, sp.dims_b, IDENT[:3],
                                           True, sp.lb), cpu_repeats)
    res["cpu_baseline"] = {"value": 2 * nprod / t / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "port",
                           "ms": round(t * 1e3, 1),
                           "sample": "the same arrays (host copies), one pinned core, median of %d: csc_no_accum "
                                     "contraction (per-column expand, in-place MSD radix, collect_duplicates)"
                                     % cpu_repeats}
    _log("contraction: %s" % res)
    return res


# The batch configuration the autotuner chose on MI355X for the batched mix
# (used instead of tuning with --no-autotune, a profiling aid).
BATCH_CONFIG = {"digit_bits": 8, "sort_block": 1024, "sort_items": 12, "small_max": 1024, "item_products": 1024,
                "window": 1536, "blocks_per_cu": 4}


def bench_batched(ctx, torch, cpu_stride=100, per_instance_runs=True, tune=True):
    """The batched mix: autotuned recorded graph vs one call per instance."""
    import ctypes

    import libnt_amd
    import oracle
    from libnt_amd import autotune, dispatch, workloads
    from libnt_amd.batch import Batch, packed_offsets
    from libnt_amd.sparse import SparseMIA

    insts = workloads.batched_contractions(10000, seed=4)
    ops = workloads.generate_operands(ctx, insts)
    if tune:
        tuned = autotune.tune(ctx, insts, ops, stride=10, passes=2, reps=2)
        _log("autotune: %s" % tuned)
    else:
        tuned = {"config": dict(BATCH_CONFIG), "ms": None, "default_ms": None}
    b = Batch(ctx, insts)
    autotune.apply_batch(ctx, b, tuned["config"])
    nprod = b.nprod(ops[0], ops[2])
    nnz = b.capture(*ops)
    _, _, c_off = b.replay()

    def replay_s():
        times = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.replay()
            times.append(time.perf_counter() - t0)
        return float(np.median(times))

    t_batched = replay_s()
    # Operand descriptors of every instance (the per-instance dispatch and the
    # per-instance accumulator choices inside the batch read them).
    a_off = packed_offsets([x.nnz_a for x in insts])
    b_off = packed_offsets([x.nnz_b for x in insts])
    ak, av, bk, bv = ops
    pairs = []
    for i, x in enumerate(insts):
        A = SparseMIA(x.dims_a, ak[a_off[i]:a_off[i + 1]], av[a_off[i]:a_off[i + 1]], ctx=ctx)
        B = SparseMIA(x.dims_b, bk[b_off[i]:b_off[i + 1]], bv[b_off[i]:b_off[i + 1]], ctx=ctx)
        pairs.append((dispatch.operand(A, x.labels_a), dispatch.operand(B, x.labels_b)))
    model_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnt_amd", "data", "dispatch_model.npy")
    model = None
    if os.path.exists(model_path):
        model = dispatch.Model(ctx, 1)
        model.load(model_path)
    # The same recorded graph with each policy's per-instance accumulators
    # (results are bit-identical; only the time differs).
    policies = {"sort": {"s": round(t_batched, 4), "dense_instances": 0, "dense_segments": 0}}
    for name in ("dense", "heuristic", "learned"):
        if name == "learned" and model is None:
            continue
        acc = dispatch.batch_accumulators(pairs, name, model)
        b.set_accumulators(acc)
        b.replay()
        policies[name] = {"s": round(replay_s(), 4),
                          "dense_instances": int(np.isin(acc, (dispatch.ACC_DENSE, dispatch.ACC_MTT)).sum()),
                          "dense_segments": b.get_param("dense_segments")}
    b.set_accumulators(dispatch.batch_accumulators(pairs, "sort"))
    b.close()
    del b
    torch.cuda.empty_cache()
    nnz_i = np.diff(c_off)
    compulsory = 16.0 * (sum(x.nnz_a + x.nnz_b for x in insts) + int(nnz))
    best = min(policies, key=lambda k: policies[k]["s"])
    t_best = policies[best]["s"]
    res = {
        "workload": "10,000 contractions: order 3-4, 1-2 contracted modes, dims log-U[8,128], nnz log-U[100,50k]",
        "products": int(nprod.sum()), "nnz_c": int(nnz),
        "batched_graph_s": round(t_batched, 4),
        "batched_gflops": round(2 * nprod.sum() / t_batched / 1e9, 2),
        "instances_per_s": round(len(insts) / t_batched, 1),
        "accumulator_policies": policies,
        "best_policy": best,
        "roofline": dict(_roofline(compulsory, t_best, "whole batched pipeline (graph replay, %s accumulators), "
                                                       "compulsory bytes 16*sum(nnzA+nnzB+nnzC)" % best),
                         # HBM bytes of one dense replay, measured by rocprofv3 on this mix.
                         traffic=_profiled_traffic("traffic_batched_10000_contractions.md",
                                                   "traffic_bytes_per_replay") if best == "dense" else None),
        "autotuned": tuned["config"],
        "autotune_subsample_ms": ({"default": round(tuned["default_ms"], 2), "tuned": round(tuned["ms"], 2)}
                                  if tune else None),
    }
    if "learned" in policies:
        res["learned_batched_s"] = policies["learned"]["s"]
        res["speedup_learned_batched_vs_batched"] = round(t_batched / policies["learned"]["s"], 3)
    if not per_instance_runs:
        if model is not None:
            model.close()
        _log("batched: %s" % res)
        return res
    # one lnt_contract_into call per instance, C into one reused buffer
    cap = int(max(nnz_i.max(), 1))
    out_k = torch.empty(cap, dtype=torch.int64, device="cuda")
    out_v = torch.empty(cap, dtype=torch.float64, device="cuda")
    L = libnt_amd.lib()

    def per_instance(strategy):
        n = ctypes.c_int64()
        shape = dispatch.ResultShape()
        used = ctypes.c_int()
        hist = np.zeros(dispatch.NUM_STRATEGIES, dtype=np.int64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i, (oa, ob) in enumerate(pairs):
            libnt_amd.check(L.lnt_contract_into(ctx.handle, ctypes.byref(oa), ctypes.byref(ob), strategy,
                                                out_k.data_ptr(), out_v.data_ptr(), cap, ctypes.byref(n),
                                                ctypes.byref(shape), ctypes.byref(used)))
            hist[used.value] += 1
        torch.cuda.synchronize()
        return time.perf_counter() - t0, hist

    per_instance(0)  # warm-up: workspaces reach their final sizes
    t_plain, h_plain = per_instance(dispatch.STRAT_AUTO)  # no model: the restated polyalgorithm
    t_fixed, _ = per_instance(0)
    res.update({
        "plain_heuristic_per_instance_s": round(t_plain, 4),
        "plain_strategy_histogram": h_plain.tolist(),
        "fixed_default_per_instance_s": round(t_fixed, 4),
        "speedup_batched_vs_plain_heuristic": round(t_plain / t_batched, 2),
    })
    if model is not None:
        model.attach()
        t_learned, h_learned = per_instance(dispatch.STRAT_AUTO)
        libnt_amd.check(L.lnt_context_set_model(ctx.handle, None))
        model.close()
        res["learned_per_instance_s"] = round(t_learned, 4)
        res["learned_strategy_histogram"] = h_learned.tolist()
        res["speedup_learned_vs_plain_heuristic"] = round(t_plain / t_learned, 3)
        res["speedup_learned_batched_vs_plain_heuristic"] = round(t_plain / policies["learned"]["s"], 2)
    del out_k, out_v
    if cpu_stride < 1:
        _log("batched: %s" % res)
        return res
    # CPU baseline: the oracle, one instance after another, on every
    # cpu_stride-th instance (operands rebuilt on the host by the same
    # counter-based generators).
    sample = list(range(0, len(insts), cpu_stride))
    host_ops = []
    for i in sample:
        x = insts[i]
        host_ops.append((oracle.gen_distinct_keys(x.nnz_a, x.dimensionality_a, workloads.KEY_SEED_A + i),
                         oracle.gen_uniform_vals(x.nnz_a, workloads.VAL_SEED_A + i),
                         oracle.gen_distinct_keys(x.nnz_b, x.dimensionality_b, workloads.KEY_SEED_B + i),
                         oracle.gen_uniform_vals(x.nnz_b, workloads.VAL_SEED_B + i)))
    with _OneCore():
        t0 = time.perf_counter()
        for (hak, hav, hbk, hbv), i in zip(host_ops, sample):
            x = insts[i]
            oracle.contract(hak, hav, x.dims_a, list(range(len(x.dims_a))), False, x.labels_a, hbk, hbv, x.dims_b,
                            list(range(len(x.dims_b))), False, x.labels_b)
        t_cpu = time.perf_counter() - t0
    sp = int(sum(nprod[i] for i in sample))
    res["cpu_baseline"] = {"value": len(sample) / t_cpu, "unit": "instances/s", "cores": 1, "kind": "port",
                           "gflops": round(2 * sp / t_cpu / 1e9, 3),
                           "sample": "every %dth instance (%d instances, %d products), one after another on one "
                                     "pinned core, %.2f s: unfold sorts + csc_no_accum + refold"
                                     % (cpu_stride, len(sample), sp, t_cpu)}
    _log("batched: %s" % res)
    return res


def bench_permutations(ctx, torch, reps=3):
    """change_sort_order of the sorted 50M-key operand A to each of the 24
    mode permutations of its order-4 dims: radix passes and time per sort, so
    the headline's one seeded permutation can be read against the spread."""
    import itertools

    ak, av, bk, bv = make_c2(ctx, torch)
    del bk, bv
    rows = []
    for perm in itertools.permutations(range(4)):
        perm = list(perm)
        t = []
        passes = 0
        for r in range(reps + 1):
            k, v = ak.clone(), av.clone()
            ctx.set_profiling(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.change_sort_order(k, v, DIMS, IDENT, perm, True)
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
            passes = ctx.profile("onesweep")[0]
            ctx.set_profiling(False)
            del k, v
        rows.append({"perm": perm, "passes": passes, "ms": round(1e3 * float(np.median(t[1:])), 4)})
    ms = np.array([r["ms"] for r in rows])
    ps = np.array([r["passes"] for r in rows])
    res = {"workload": "change_sort_order of 50M sorted keys (256^4) to every permutation of 4 modes",
           "per_permutation": rows, "passes_mean": round(float(ps.mean()), 3),
           "ms_min": float(ms.min()), "ms_median": float(np.median(ms)), "ms_mean": round(float(ms.mean()), 4),
           "ms_max": float(ms.max()), "headline_perm": _perm(),
           "headline_perm_ms": next(r["ms"] for r in rows if r["perm"] == _perm())}
    del ak, av
    torch.cuda.empty_cache()
    _log("permutations: %s" % {k: v for k, v in res.items() if k != "per_permutation"})
    return res


def bench_cpp_front_end(steps=10, warmup=2, host_steps=2):
    """The sort/merge step written against the C++ front end
    (tests/cpp/front_end_step.cpp), with the tensors resident in HBM and with
    the reference's host vectors; compiled with g++ here, run as a child
    process (its own context on GPU 0)."""
    import subprocess
    import tempfile

    root = os.path.dirname(os.path.abspath(__file__))
    src = os.path.join(root, "tests", "cpp", "front_end_step.cpp")
    libdir = os.path.join(root, "libnt_amd")
    res = {"workload": "the sort/merge-at-scale step through include/lnt/SparseMIA.hpp: A.sort(perm), B.sort(perm), "
                       "C(i,j,k,l)=A(i,j,k,l)+B(i,j,k,l), D(i,j,k,l)=A(!i,!j,!k,!l)*B(!i,!j,!k,!l)"}
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "front_end_step")
        r = subprocess.run(["g++", "-std=c++17", "-O2", src, "-I", os.path.join(root, "include"), "-L", libdir,
                            "-llnt", "-Wl,-rpath," + libdir, "-o", exe], capture_output=True, text=True)
        if r.returncode != 0:
            return {"error": "g++ failed: " + r.stderr[-400:]}
        perm = [str(x) for x in _perm()]
        for storage, n, w in (("device", steps, warmup), ("host", host_steps, 1)):
            out = subprocess.run([exe, storage, str(n), str(w)] + perm, capture_output=True, text=True, timeout=600)
            if out.returncode != 0:
                res[storage] = {"error": (out.stdout + out.stderr)[-400:]}
                continue
            res[storage] = json.loads(out.stdout.strip().splitlines()[-1])
    _log("C++ front end: %s" % res)
    return res


def bench_learned(ctx):
    """Held-out regret of the shipped model (committed MI355X timings)."""
    from libnt_amd import dispatch

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnt_amd", "data")
    tpath, mpath = os.path.join(here, "dispatch_timings.npz"), os.path.join(here, "dispatch_model.npy")
    if not (os.path.exists(tpath) and os.path.exists(mpath)):
        return None
    d = np.load(tpath, allow_pickle=False)
    m = dispatch.Model(ctx, 1)
    m.load(mpath)
    ms = np.exp2(d["T_test"].astype(np.float64))
    learned = dispatch.choose(m.predict(d["X_test"]))
    idx = np.arange(len(ms))
    res = {"held_out": int(len(ms)), "train": int(len(d["X_train"]))}
    for name, ch in (("learned", learned), ("heuristic", d["H_test"]), ("fixed_default", np.zeros(len(ms), int))):
        r = dispatch.regret(ms, ch)
        res[name] = {"total_ms": round(float(ms[idx, ch].sum()), 2), "mean_regret": round(float(r.mean()), 4),
                     "median_regret": round(float(np.median(r)), 4)}
    res["best_total_ms"] = round(float(ms.min(axis=1).sum()), 2)
    m.close()
    return res


def bench_host_buffers(ctx, torch, ak, av, bk, bv, reps=4):
    """The headline step on host arrays: each call copies its operands to HBM
    and its results back (LNT_MEM_HOST), as a caller holding the reference's
    host vectors would use the C-ABI."""
    perm = _perm()
    orders = [IDENT, perm]
    n = ak.shape[0] + bk.shape[0]
    res = {}
    for mode in ("pinned", "pageable"):
        if mode == "pinned":
            h = [t.cpu().pin_memory() for t in (ak, av, bk, bv)]
            outs = [torch.empty(n, dtype=dt).pin_memory() for dt in (torch.int64, torch.float64)]
            m = min(ak.shape[0], bk.shape[0])
            outs += [torch.empty(m, dtype=dt).pin_memory() for dt in (torch.int64, torch.float64)]
        else:
            h = [t.cpu().numpy().copy() for t in (ak, av, bk, bv)]
            outs = [np.empty(n, dtype=dt) for dt in (np.int64, np.float64)]
            m = min(ak.shape[0], bk.shape[0])
            outs += [np.empty(m, dtype=dt) for dt in (np.int64, np.float64)]

        def step(s):
            cur, new = orders[s % 2], orders[(s + 1) % 2]
            ctx.change_sort_order(h[0], h[1], DIMS, cur, new, True)
            ctx.change_sort_order(h[2], h[3], DIMS, cur, new, True)
            ctx.merge_into(h[0], h[1], h[2], h[3], 0, outs[0], outs[1])
            ctx.merge_into(h[0], h[1], h[2], h[3], 2, outs[2], outs[3])

        step(0)
        step(1)
        t = time.perf_counter()
        for s in range(reps):
            step(s)
        el = (time.perf_counter() - t) / reps
        res[mode] = {"ms_per_step": round(el * 1e3, 1), "keys_per_s": n / el}
        _log("host buffers (%s): %s" % (mode, res[mode]))
        del h, outs
    return res


SECTIONS = ("headline", "unsorted_sort", "permutations", "small_contraction", "contraction", "batched", "learned",
            "cpp_front_end")


def bench_headline(ctx, torch, args):
    """The sort/merge-at-scale step (the `value` of the JSON line)."""
    import libnt_amd

    ak, av, bk, bv = make_c2(ctx, torch)
    host = None if args.no_cpu_baseline or args.cpu_scale > 1 else [t.cpu().numpy() for t in (ak, av, bk, bv)]
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
            "radix_passes_per_sort": launches / (2 * args.steps),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "onesweep radix pass (8-bit digit, %dx%d tile)" % (ctx.get_param("sort_block"),
                                                                      ctx.get_param("sort_items")),
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
    if args.host_buffers:
        out["host_buffers"] = bench_host_buffers(ctx, torch, ak, av, bk, bv)
    del ak, av, bk, bv, cu_k, cu_v, ch_k, ch_v
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(*(host or (None,) * 4), scale=args.cpu_scale)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-scale", type=int, default=1,
                    help="1: the CPU baseline runs on host copies of the GPU's inputs; s > 1: on the generator's "
                         "first 1/s of the workload")
    ap.add_argument("--no-extras", action="store_true", help="only the sort/merge workload")
    ap.add_argument("--section", choices=("all",) + SECTIONS, default="all",
                    help="run one section only (a profiling aid: the line then holds that section alone)")
    ap.add_argument("--graph-only", action="store_true",
                    help="batched section: the recorded graph only, no per-instance loops (a profiling aid)")
    ap.add_argument("--strategies", type=lambda x: int(x, 0), default=None,
                    help="contraction section: time only the strategies of this bit mask (a profiling aid)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched section: the stored MI355X configuration instead of tuning (a profiling aid)")
    ap.add_argument("--calibrate", action="store_true",
                    help="after the timed steps, run a plain 1.6 GB write (fill) and read (sum) so that "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE can be scaled by kernels of known traffic")
    ap.add_argument("--host-buffers", action="store_true",
                    help="also time the step on host arrays (PCIe copies included; not the headline value)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import torch

    import libnt_amd

    stream = torch.cuda.current_stream()
    ctx = libnt_amd.Context(0, stream=ctypes.c_void_p(stream.cuda_stream))
    want = set(SECTIONS) if args.section == "all" else {args.section}
    if args.no_extras:
        want &= {"headline"}
    out = bench_headline(ctx, torch, args) if "headline" in want else {"section": args.section}
    if "unsorted_sort" in want:
        out["unsorted_sort"] = bench_unsorted_sort(ctx, torch)
        _log("unsorted sort: %s" % (out["unsorted_sort"],))
    if "permutations" in want:
        out["permutations"] = bench_permutations(ctx, torch)
    if "small_contraction" in want:
        out["small_contraction"] = bench_small_contraction(ctx, torch)
    if "contraction" in want:
        out["contraction"] = bench_contraction(ctx, torch, cpu_repeats=0 if args.no_cpu_baseline else 3,
                                               mask=args.strategies)
    if "batched" in want:
        out["batched"] = bench_batched(ctx, torch, cpu_stride=0 if args.no_cpu_baseline else 100,
                                       per_instance_runs=not args.graph_only, tune=not args.no_autotune)
    if "learned" in want:
        out["learned"] = bench_learned(ctx)
    if "cpp_front_end" in want:
        out["cpp_front_end"] = bench_cpp_front_end()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
