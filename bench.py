#!/usr/bin/env python3
"""Throughput of meshoid's neighbour-weighted operator path on one MI355X.

Workload: 2^24 clustered particles (256 Gaussian clumps, sigma log-uniform in
[0.05, 5], clump masses dN/dM ~ M^-2, 10% uniform background, 100-unit box,
float32 positions), des_ngb = 32, one float32 field f = sin(kx) sin(ky) sin(kz)
with k = 2 pi / 10.  One step = the whole path on device-resident inputs:
spatial index build (Morton radix sort + tree), exact kNN + smoothing length +
density, and the least-squares gradient of f.

Prints ONE JSON line (the metric contract), with
  roofline      the dominant kernel's algorithmic bytes / its mean duration,
                measured with HIP events on the library's stream
  cpu_baseline  the reference algorithm on the host cores (scipy cKDTree --
                the reference's own kNN call -- plus the FP64 oracle's HsmlIter
                and gradient weights) on a 2^21-particle sample of the workload
  accuracy      h / rho / grad f errors against the FP64 oracle on a sample

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "particles/s (hsml+ρ+∇f) at 16.7M clustered, ngb=32; learned vs plain GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(region: str, k: int, K: int) -> float:
    """Algorithmic bytes per particle of each launch region: own inputs + own
    outputs + one read of every (particle, neighbour) pair."""
    if region == "knn_pass1":  # own pos + k neighbour positions + h, rho
        return 12 + 12 * k + 8
    if region == "knn_pass2":  # same gathers again: neighbour list collection
        return 12 + 12 * k
    if region in ("apply_D1", "apply_D2"):  # own pos, h, field; k x (pos, field); grad
        return 12 + 4 + 4 * K + k * (12 + 4 * K) + 12 * K
    return 0.0


def cpu_baseline(n, k, seed):
    """Reference algorithm on the host: cKDTree.query (Meshoid.py:145,157),
    HsmlIter bisection (kernel_density.py:14-57), gradient_weights
    (derivatives.py:117-138) and the einsum of Meshoid.D (:258,268)."""
    from scipy.spatial import cKDTree

    from meshoid_amd import synthetic
    from oracle import oracle as orc

    pos = synthetic.clustered(n, seed=seed, nclump=256, background=0.1).astype(np.float64)
    f, _ = synthetic.sin3(pos)
    t0 = time.perf_counter()
    tree = cKDTree(pos)
    d, ngb = tree.query(pos, k, workers=-1)
    h = orc.hsml_iter(d, error_norm=1e-13)
    G = orc.gradient_weights(pos, ngb, h, np.arange(n))
    orc.apply_weights(G, ngb, np.arange(n), f)
    dt = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"value": n / dt, "unit": "particles/s", "cores": int(cores), "kind": "port",
            "sample": f"{n} clustered particles (same generator, seed {seed}), k={k}, K=1: "
                      f"scipy cKDTree(workers=-1) + FP64 OpenMP oracle HsmlIter/gradient_weights/"
                      f"einsum ({orc.num_threads()} threads), {dt:.2f} s"}


def accuracy(pos, f, h, rho, d1, k, nsample, seed):
    """Errors of the GPU result against the FP64 oracle on a random sample."""
    from oracle import oracle as orc

    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(pos), nsample, replace=False))
    p64 = pos.astype(np.float64)
    d, ngb = orc.knn(p64, k, queries=p64[idx])
    h_o = orc.hsml_iter(d, error_norm=1e-13)
    G = orc.gradient_weights(p64, ngb, h_o, idx)
    g_o = orc.apply_weights(G, ngb, idx, f.astype(np.float64))
    rho_o = k / (4 * np.pi / 3 * h_o**3)
    eh = np.abs(h[idx] / h_o - 1)
    er = np.abs(rho[idx] / rho_o - 1)
    eg = np.linalg.norm(d1[idx] - g_o, axis=1) / np.maximum(np.linalg.norm(g_o, axis=1), 1e-300)
    return {"sample": int(nsample), "h_max_rel": float(eh.max()), "h_median_rel": float(np.median(eh)),
            "rho_max_rel": float(er.max()), "gradf_max_rel": float(eg.max()),
            "gradf_p99999_rel": float(np.quantile(eg, 0.99999)),
            "gradf_median_rel": float(np.median(eg))}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=1 << 21)
    ap.add_argument("--check", type=int, default=1 << 16, help="oracle sample size (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stats", action="store_true", help="kNN work counters (slower)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus}, WORLD_SIZE {world} requested)", file=sys.stderr)
        sys.exit(2)

    import torch

    import meshoid_amd
    from meshoid_amd import engine, synthetic

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n, k, K = args.n, args.k, 1
    log(f"generating {n} clustered particles")
    pos = synthetic.clustered(n, seed=args.seed, nclump=256, background=0.1)
    f, _ = synthetic.sin3(pos)
    f = f.astype(np.float32)
    pos_d = torch.from_numpy(pos).to(dev)
    f_d = torch.from_numpy(f).to(dev)
    out = {name: torch.empty(n, dtype=torch.float32, device=dev) for name in ("h", "rho", "vol")}
    out["status"] = torch.empty(n, dtype=torch.int32, device=dev)
    d1 = torch.empty((n, K, 3), dtype=torch.float32, device=dev)
    st = torch.empty(n, dtype=torch.int32, device=dev)
    ix = engine.SpatialIndex()
    nl = engine.NeighborList()

    def step():
        ix.build(pos_d)
        engine.knn_hsml(ix, k, keep_dist=False, nlist=nl, out=out)
        engine.apply_D(nl, f_d, out1=d1, status=st)

    log("warmup")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    engine.prof_read()
    engine.prof_enable(2 if args.stats else 1)
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    prof = engine.prof_read()
    engine.prof_enable(0)
    stats = {nm: cnt for nm, (cnt, _) in prof.items() if nm.startswith("stat_")}
    kernels = {name: tot / cnt for name, (cnt, tot) in prof.items() if not name.startswith("stat_")}
    if stats:
        w = max(stats.get("stat_waves", 1), 1)
        log("per-wave work:", {a[5:]: round(b / w, 2) for a, b in stats.items()})
    log("per-step region ms:", {a: round(b, 3) for a, b in kernels.items()})
    dom = max((r for r in kernels if alg_bytes(r, k, K) > 0), key=lambda r: kernels[r])
    ach = alg_bytes(dom, k, K) * n / (kernels[dom] * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"{dom}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    status = st.cpu().numpy() | out["status"].cpu().numpy()
    acc = None
    if args.check:
        log("oracle accuracy sample")
        acc = accuracy(pos, f, out["h"].cpu().numpy().astype(np.float64),
                       out["rho"].cpu().numpy().astype(np.float64),
                       d1[:, 0, :].cpu().numpy().astype(np.float64), k, args.check, args.seed + 1)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(args.cpu_n, k, args.seed)
    rec = {
        "metric": METRIC,
        "value": n / dt,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "clustered_2^24_k32_K1", "n": n, "des_ngb": k, "fields": K,
                   "form": "plain", "positions": "float32", "generator_seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": ach / HBM_PEAK, "traffic": traffic,
                     "alg_bytes_per_particle": alg_bytes(dom, k, K)},
        "pipeline_roofline_frac": n * 552 / dt / HBM_PEAK,
        "kernels_ms": {a: round(b, 4) for a, b in kernels.items()},
        "status_counts": {"ambiguous_fp64_resolved": int(((status & 1) != 0).sum()),
                          "rank_deficient": int(((status & 64) != 0).sum()),
                          "hsml_noconv": int(((status & 4) != 0).sum())},
        "accuracy": acc,
        "cpu_baseline": cpu,
        "library": meshoid_amd._lib.LIB_PATH,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
