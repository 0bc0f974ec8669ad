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

Further objects on the same line (skip with --no-extras):
  contraction: the large skewed contraction A(i,j,k)*B(k,l,m), A 4096x4096x512
    (858,993 nnz), B 512x4096x64 (13,422 nnz), i and l Zipf(1.1): whole
    contraction (unfold sorts + lattice product + refold) per accumulator,
    flops/s = 2 * scalar products / time;
  batched: 10,000 independent contractions (order 3-4, 1-2 contracted modes,
    dims log-U[8,128], nnz log-U[100,50k]) as one recorded hipGraph with
    autotuned parameters, against one lnt_contract call per instance with the
    restated polyalgorithm's strategy and with the learned model's strategy;
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


def cpu_baseline(scale=8, repeats=5):
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
            "cpu_model": _cpu_model(),
            "nproc": len(old),
            "sample": "1/%d-scale c2 step (%d+%d keys) on one pinned core: 2x change_sort_order (relinearise + "
                      "in-place MSD radix 2048/11-bit, thresholds 3000/50) + scan merge A+B + Hadamard (the "
                      "reference's reserve(nA*nB/2^Inter) replaced by min(nA, nB)); median of %d, %.2f s each"
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


def bench_contraction(ctx, torch):
    """Large skewed contraction, both accumulators, timed whole."""
    from libnt_amd import dispatch

    sp = dispatch.Spec("skewed", [4096, 4096, 512], "ijk", [512, 4096, 64], "klm", 858993, 13422,
                       "default", "default", (0,), (1,))
    A, B = dispatch.materialize(ctx, sp, 31)
    # scalar products: sum over k of nnzA(:,k) * nnzB(k,:)
    ka = A.keys.cpu().numpy()
    kb = B.keys.cpu().numpy()
    ca = np.bincount(ka // (4096 * 4096), minlength=512)
    cb = np.bincount(kb % 512, minlength=512)
    nprod = int((ca.astype(np.int64) * cb).sum())
    ms = dispatch.time_strategies(ctx, A, sp.la, B, sp.lb, 0b0111, 3)
    res = {"workload": "A(i,j,k)*B(k,l,m), A 4096x4096x512 nnz 858993, B 512x4096x64 nnz 13422, "
                       "i and l Zipf(1.1); unfold sorts + lattice product + refold",
           "products": nprod, "unit": "GFLOP/s (2 flops per scalar product)"}
    for s in (0, 1, 2):
        res[dispatch.strategy_name(s)] = {"ms": round(ms[s], 3), "gflops": round(2 * nprod / (ms[s] * 1e-3) / 1e9, 2)}
    best = min((0, 1, 2), key=lambda s: ms[s])
    res["best"] = dispatch.strategy_name(best)
    _log("contraction: %s" % res)
    return res


def bench_batched(ctx, torch):
    """The batched mix: autotuned recorded graph vs one call per instance."""
    import ctypes

    import libnt_amd
    from libnt_amd import autotune, dispatch, workloads
    from libnt_amd.batch import Batch, packed_offsets
    from libnt_amd.sparse import SparseMIA

    insts = workloads.batched_contractions(10000, seed=4)
    ops = workloads.generate_operands(ctx, insts)
    tuned = autotune.tune(ctx, insts, ops, stride=10, passes=2, reps=2)
    _log("autotune: %s" % tuned)
    b = Batch(ctx, insts)
    autotune.apply_batch(ctx, b, tuned["config"])
    nprod = b.nprod(ops[0], ops[2])
    nnz = b.capture(*ops)
    b.replay()
    times = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.replay()
        times.append(time.perf_counter() - t0)
    t_batched = float(np.median(times))
    b.close()
    del b
    torch.cuda.empty_cache()
    # one lnt_contract call per instance (library-allocated result, freed)
    a_off = packed_offsets([x.nnz_a for x in insts])
    b_off = packed_offsets([x.nnz_b for x in insts])
    ak, av, bk, bv = ops
    pairs = []
    for i, x in enumerate(insts):
        A = SparseMIA(x.dims_a, ak[a_off[i]:a_off[i + 1]], av[a_off[i]:a_off[i + 1]], ctx=ctx)
        B = SparseMIA(x.dims_b, bk[b_off[i]:b_off[i + 1]], bv[b_off[i]:b_off[i + 1]], ctx=ctx)
        pairs.append((dispatch.operand(A, x.labels_a), dispatch.operand(B, x.labels_b)))
    L = libnt_amd.lib()

    def per_instance(strategy):
        ck, cv = ctypes.c_void_p(), ctypes.c_void_p()
        n = ctypes.c_int64()
        shape = dispatch.ResultShape()
        used = ctypes.c_int()
        hist = np.zeros(dispatch.NUM_STRATEGIES, dtype=np.int64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for oa, ob in pairs:
            libnt_amd.check(L.lnt_contract(ctx.handle, ctypes.byref(oa), ctypes.byref(ob), strategy, ctypes.byref(ck),
                                           ctypes.byref(cv), ctypes.byref(n), ctypes.byref(shape), ctypes.byref(used)))
            L.lnt_free(ctx.handle, ck, libnt_amd.MEM_DEVICE)
            L.lnt_free(ctx.handle, cv, libnt_amd.MEM_DEVICE)
            hist[used.value] += 1
        torch.cuda.synchronize()
        return time.perf_counter() - t0, hist

    t_plain, h_plain = per_instance(dispatch.STRAT_AUTO)  # no model: the restated polyalgorithm
    res = {
        "workload": "10,000 contractions: order 3-4, 1-2 contracted modes, dims log-U[8,128], nnz log-U[100,50k]",
        "products": int(nprod.sum()), "nnz_c": int(nnz),
        "batched_graph_s": round(t_batched, 4),
        "batched_gflops": round(2 * nprod.sum() / t_batched / 1e9, 2),
        "instances_per_s": round(len(insts) / t_batched, 1),
        "autotuned": tuned["config"], "autotune_subsample_ms": {"default": round(tuned["default_ms"], 2),
                                                                 "tuned": round(tuned["ms"], 2)},
        "plain_heuristic_per_instance_s": round(t_plain, 4),
        "plain_strategy_histogram": h_plain.tolist(),
        "speedup_batched_vs_plain_heuristic": round(t_plain / t_batched, 2),
    }
    model_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnt_amd", "data", "dispatch_model.npy")
    if os.path.exists(model_path):
        m = dispatch.Model(ctx, 1)
        m.load(model_path)
        m.attach()
        t_learned, h_learned = per_instance(dispatch.STRAT_AUTO)
        libnt_amd.check(L.lnt_context_set_model(ctx.handle, None))
        m.close()
        res["learned_per_instance_s"] = round(t_learned, 4)
        res["learned_strategy_histogram"] = h_learned.tolist()
        res["speedup_learned_vs_plain_heuristic"] = round(t_plain / t_learned, 3)
    _log("batched: %s" % res)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-scale", type=int, default=8)
    ap.add_argument("--no-extras", action="store_true", help="only the sort/merge workload")
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
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scale=args.cpu_scale)
    if not args.no_extras:
        del ak, av, bk, bv, cu_k, cu_v, ch_k, ch_v
        torch.cuda.empty_cache()
        out["unsorted_sort"] = bench_unsorted_sort(ctx, torch)
        _log("unsorted sort: %s" % (out["unsorted_sort"],))
        out["contraction"] = bench_contraction(ctx, torch)
        out["batched"] = bench_batched(ctx, torch)
        out["learned"] = bench_learned(ctx)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
