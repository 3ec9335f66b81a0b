"""Writes the kNN golden vectors under tests/golden/.

The reference obtains neighbours from ``scipy.spatial.cKDTree(pos, boxsize)``
and ``.query(x, des_ngb)`` (src/meshoid/Meshoid.py:145,157).  scipy is a
third-party dependency that the reference does not pin (requirements.txt);
these fixtures were produced with the scipy in this container (recorded in each
file as ``scipy_version``).  Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
from scipy.spatial import cKDTree

HERE = os.path.dirname(os.path.abspath(__file__))


def clustered(n, seed, box=100.0, nclump=16):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, box, (nclump, 3))
    sig = np.exp(rng.uniform(np.log(0.05), np.log(5.0), nclump))
    w = rng.pareto(1.0, nclump) + 1.0
    counts = rng.multinomial(n, w / w.sum())
    pts = [c + s * rng.standard_normal((m, 3)) for c, s, m in zip(centers, sig, counts)]
    return np.concatenate(pts).astype(np.float32)


def case(name, pos, k, boxsize=None, nq=None, seed=0):
    rng = np.random.default_rng(seed)
    n = pos.shape[0]
    qidx = np.arange(n) if nq is None else np.sort(rng.choice(n, nq, replace=False))
    tree = cKDTree(np.asarray(pos, dtype=np.float64), boxsize=boxsize)
    dist, ngb = tree.query(np.asarray(pos, dtype=np.float64)[qidx], k)
    np.savez_compressed(
        os.path.join(HERE, f"knn_{name}.npz"), pos=pos, qidx=qidx.astype(np.int64),
        ngb=ngb.astype(np.int32), dist=dist, k=k,
        boxsize=-1.0 if boxsize is None else float(boxsize), scipy_version=scipy.__version__)
    print(name, pos.shape, pos.dtype, "k", k, "nq", len(qidx))


if __name__ == "__main__":
    r = np.random.default_rng(1)
    case("uniform3d_f64", r.random((1024, 3)), 32)
    case("periodic3d_f64", np.random.default_rng(2).random((1024, 3)), 32, boxsize=1.0)
    case("clustered3d_f32", clustered(4096, 3), 32, nq=512, seed=3)
    case("uniform2d_f64", np.random.default_rng(4).random((1024, 2)), 20)
    case("periodic3d_k100", np.random.default_rng(42).random((4096, 3)), 100, boxsize=1.0,
         nq=128, seed=5)
