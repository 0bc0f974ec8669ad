"""Measures the dispatch training and held-out sets on the GPU and trains the
shipped cost model (run on one MI355X from the repository root).  A
maintainer's script that writes this data directory; the directory is not a
Python package and nothing in the library imports this file:

    python libnt_amd/data/make_dispatch_data.py --part train --start 0 --count 7000 --out DIR
    python libnt_amd/data/make_dispatch_data.py --part train --start 7000 --count 7000 --out DIR
    python libnt_amd/data/make_dispatch_data.py --part train --start 14000 --count 6000 --out DIR
    python libnt_amd/data/make_dispatch_data.py --part test --start 0 --count 2000 --out DIR
    python libnt_amd/data/make_dispatch_data.py --part trainbatch --start 0 --count 10000 --out DIR
    python libnt_amd/data/make_dispatch_data.py --part testbatch --start 0 --count 2000 --out DIR
    python libnt_amd/data/make_dispatch_data.py --part fit --out DIR

(chunks keep each run short).  Writes DIR/dispatch_timings.npz (features, log2-ms timings of all 20
strategies, the restated polyalgorithm's choice) and DIR/dispatch_model.npy
(the model trained on the training split by the library's GPU kernels).
Instances come from ``dispatch.mixed_specs`` (train seed 1, held-out seed 2)
and ``dispatch.batch_specs`` (the batched contraction mix with unsorted
operands: train seed 5, held-out seed 6; the benchmark's own mix is seed 4).
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

EPOCHS, LR = 300, 1e-3


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--part", required=True, choices=["train", "test", "trainbatch", "testbatch", "fit"])
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--count", type=int, default=0)
    ap.add_argument("--max-nnz", type=int, default=200000)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    args = ap.parse_args(argv)

    import glob

    import torch

    import libnt_amd
    from libnt_amd import dispatch

    ctx = libnt_amd.Context(0, torch.cuda.current_stream().cuda_stream)
    t0 = time.time()

    def log(msg):
        print("[%6.1fs] %s" % (time.time() - t0, msg), flush=True)

    os.makedirs(args.out, exist_ok=True)
    seeds = {"train": 1, "test": 2, "trainbatch": 5, "testbatch": 6}
    if args.part in seeds:
        seed = seeds[args.part]
        if args.part.endswith("batch"):
            specs = dispatch.batch_specs(args.start + args.count, seed)[args.start:]
            name = "dispatch_%s_b%06d.npz" % (args.part[: -len("batch")], args.start)
        else:
            specs = dispatch.mixed_specs(args.start + args.count, seed, args.max_nnz)[args.start:]
            name = "dispatch_%s_%06d.npz" % (args.part, args.start)
        X, T, H = dispatch.collect_timings(ctx, specs, seed * 1000003 + args.start, args.reps, log)
        path = os.path.join(args.out, name)
        np.savez_compressed(path, X=X, T=T, H=H)
        log("%s instances %d..%d -> %s" % (args.part, args.start, args.start + args.count, path))
        return 0

    def load(name):
        parts = sorted(glob.glob(os.path.join(args.out, "dispatch_%s_*.npz" % name)))
        if not parts:
            raise SystemExit("no %s parts under %s" % (name, args.out))
        ds = [np.load(p, allow_pickle=False) for p in parts]
        return (np.concatenate([d["X"] for d in ds]), np.concatenate([d["T"] for d in ds]),
                np.concatenate([d["H"] for d in ds]))

    Xtr, Ttr, Htr = load("train")
    Xte, Tte, Hte = load("test")
    model = dispatch.Model(ctx, 1)
    loss = model.train(Xtr, Ttr, epochs=EPOCHS, batch=256, lr=LR, seed=7)
    log("trained on %d instances: loss %.4f" % (len(Xtr), loss))
    ms = np.exp2(Tte.astype(np.float64))
    learned = dispatch.choose(model.predict(Xte))
    for name, ch in (("learned", learned), ("heuristic", Hte), ("fixed 0", np.zeros(len(ms), int))):
        r = dispatch.regret(ms, ch)
        log("%-9s total %.1f ms  mean regret %.3f  median %.3f" % (name, ms[np.arange(len(ms)), ch].sum(), r.mean(),
                                                                    np.median(r)))
    log("best      total %.1f ms over %d held-out instances" % (ms.min(axis=1).sum(), len(ms)))
    np.savez_compressed(os.path.join(args.out, "dispatch_timings.npz"), X_train=Xtr, T_train=Ttr, H_train=Htr,
                        X_test=Xte, T_test=Tte, H_test=Hte, epochs=EPOCHS, lr=LR)
    model.save(os.path.join(args.out, "dispatch_model.npy"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
