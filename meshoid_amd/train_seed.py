"""Trains the learned smoothing-length seed on the GPU and saves it.

    python -m meshoid_amd.train_seed [--nsets 32] [--n 1048576] [--k 32]

Training data: the plain method's smoothing lengths on `nsets` seeded
clustered sets of n particles (256 Gaussian clumps, 10% background).  The
forward/backward passes and Adam run in libmeshoid_hip.so.  A held-out set
gives the validation error of log h.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import torch

from . import engine, seed, synthetic


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--nsets", type=int, default=32)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    F, T = seed.training_set(args.nsets, args.n, args.k, seed0=1000, log=log)
    t1 = time.perf_counter()
    model, hist = seed.train(F, T, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    t2 = time.perf_counter()
    # held-out validation set
    pos = torch.from_numpy(synthetic.clustered(args.n, seed=5000, nclump=256,
                                               background=0.1)).cuda()
    ix = engine.SpatialIndex().build(pos)
    h = engine.knn_hsml(ix, args.k, keep_dist=False)["h"]
    e = torch.log(model.predict(ix)) - torch.log(h)
    q = torch.quantile(e.abs()[: 1 << 20].float(), torch.tensor([0.5, 0.9, 0.99], device=e.device))
    model.meta.update({"nsets": args.nsets, "n": args.n, "k": args.k,
                       "val_rms_log_err": float(e.pow(2).mean().sqrt()),
                       "val_abs_log_err_q50_q90_q99": [float(x) for x in q],
                       "data_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)})
    path = args.out or seed.default_path(args.k)
    model.save(path)
    print(json.dumps({"saved": path, **model.meta}), flush=True)


if __name__ == "__main__":
    main()
