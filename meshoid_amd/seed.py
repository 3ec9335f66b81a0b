"""Learned initial guess for the smoothing length.

A 28-64-64-1 SiLU network maps coarse-cell counts around a particle (the 3x3x3
block of cells of 1/128 of the index frame, as log(1 + count), plus log(cell
size)) to log(h / cell).  Inference (BF16 MFMA), training (forward + backward
+ Adam) and the features all run in libmeshoid_hip.so; PyTorch only holds the
device buffers and draws the minibatch permutation.

The prediction seeds two things in ``engine.knn_hsml``:
  * ``seed_h``: the starting point of the smoothing-length Newton solve;
  * ``seed_r = beta * h_pred``: an initial search radius for the exact kNN.
Rows where either guess is unusable fall back to the plain method, so results
never depend on the network.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import engine, synthetic
from ._lib import lib, raise_for

NIN, NH, NPARAM = 28, 64, 6081
HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")


def default_path(k=32):
    return os.path.join(DATA, f"seed_mlp_k{k}.npz")


def _init_params(rng):
    """He-style initialisation in the blob order W1, b1, W2, b2, w3, b3."""
    W1 = rng.standard_normal((NH, NIN)) * np.sqrt(2.0 / NIN)
    W2 = rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH)
    w3 = rng.standard_normal(NH) * np.sqrt(1.0 / NH)
    return np.concatenate([W1.ravel(), np.zeros(NH), W2.ravel(), np.zeros(NH), w3, [0.0]]
                          ).astype(np.float32)


class SeedModel:
    """Network parameters (float32 master copy), input normalisation and the
    packed BF16 inference blob, all on one device."""

    def __init__(self, P, mu, sd, device=None, meta=None):
        dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.P = torch.as_tensor(np.asarray(P, np.float32)).to(dev).contiguous()
        self.mu = torch.as_tensor(np.asarray(mu, np.float32)).to(dev).contiguous()
        self.sd = torch.as_tensor(np.asarray(sd, np.float32)).to(dev).contiguous()
        self.meta = dict(meta or {})
        self._blob = None

    @property
    def device(self):
        return self.P.device

    def blob(self):
        if self._blob is None:
            self._blob = torch.empty(int(lib.mhip_mlp_blob_bytes()), dtype=torch.uint8,
                                     device=self.device)
            ctx = engine.context(self.device)
            raise_for(lib.mhip_mlp_pack(ctx, engine.ptr(self.P), engine.ptr(self.mu),
                                        engine.ptr(self.sd), engine.ptr(self._blob)), ctx)
        return self._blob

    def predict(self, index: engine.SpatialIndex) -> torch.Tensor:
        """Predicted smoothing length per particle (original row order)."""
        h = torch.empty(index.n, dtype=torch.float32, device=self.device)
        ctx = engine.context(self.device)
        raise_for(lib.mhip_seed_predict(ctx, index.h, engine.ptr(self.blob()), engine.ptr(h)), ctx)
        return h

    def save(self, path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, P=self.P.cpu().numpy(), mu=self.mu.cpu().numpy(),
                 sd=self.sd.cpu().numpy(), meta=np.array(repr(self.meta)))

    @classmethod
    def load(cls, path, device=None):
        z = np.load(path, allow_pickle=False)
        meta = {}
        if "meta" in z:
            import ast

            meta = ast.literal_eval(str(z["meta"]))
        return cls(z["P"], z["mu"], z["sd"], device=device, meta=meta)


def load_default(k=32, device=None):
    path = default_path(k)
    return SeedModel.load(path, device) if os.path.exists(path) else None


def features(index: engine.SpatialIndex) -> torch.Tensor:
    """Raw features (n, 28) in original row order."""
    F = torch.empty((index.n, NIN), dtype=torch.float32, device=index.pos.device)
    ctx = engine.context(index.pos.device)
    raise_for(lib.mhip_seed_features(ctx, index.h, engine.ptr(F)), ctx)
    return F


def training_set(nsets=32, n=1 << 20, k=32, seed0=1000, device=None, log=None):
    """(F, T): features and targets log(h / cell) from the plain method on
    `nsets` seeded clustered sets of n particles (256 clumps, 10% background)."""
    dev = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    Fs, Ts = [], []
    ix = engine.SpatialIndex()
    nl = engine.NeighborList()
    for s in range(nsets):
        pos = torch.from_numpy(synthetic.clustered(n, seed=seed0 + s, nclump=256,
                                                   background=0.1)).to(dev)
        ix.build(pos)
        out = engine.knn_hsml(ix, k, keep_dist=False, nlist=nl)
        F = features(ix)
        T = torch.log(out["h"]) - F[:, 27]
        Fs.append(F)
        Ts.append(T)
        if log:
            log(f"training set {s + 1}/{nsets}: {n} particles")
    return torch.cat(Fs), torch.cat(Ts).contiguous()


def train(F, T, epochs=2, batch=8192, lr=2e-3, seed=0, log=None, mu=None, sd=None):
    """Adam on the mean squared error of log(h / cell), all in the library's
    kernels.  Returns (SeedModel, per-epoch mean losses)."""
    dev = F.device
    if mu is None:
        mu = F.mean(0)
        sd = F.std(0).clamp_min(1e-3)
    rng = np.random.default_rng(seed)
    P = torch.from_numpy(_init_params(rng)).to(dev)
    m = torch.zeros_like(P)
    v = torch.zeros_like(P)
    grad = torch.zeros_like(P)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    ctx = engine.context(dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    step, hist = 0, []
    ntot = F.shape[0]
    for ep in range(epochs):
        perm = torch.randperm(ntot, device=dev, generator=gen)
        loss.zero_()
        nb_done = 0
        for b0 in range(0, ntot - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            Fb = F[idx].contiguous()
            Tb = T[idx].contiguous()
            step += 1
            rc = lib.mhip_mlp_train_step(ctx, engine.ptr(Fb), engine.ptr(Tb), batch,
                                         engine.ptr(mu), engine.ptr(sd), engine.ptr(P),
                                         engine.ptr(m), engine.ptr(v), engine.ptr(grad),
                                         engine.ptr(loss), step, float(lr))
            raise_for(rc, ctx)
            nb_done += 1
        hist.append(float(loss.item()) / max(nb_done, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean batch loss {hist[-1]:.4f}")
    model = SeedModel(P.cpu().numpy(), mu.cpu().numpy(), sd.cpu().numpy(), device=dev,
                      meta={"epochs": epochs, "batch": batch, "lr": lr, "ntrain": int(ntot),
                            "losses": hist})
    return model, hist


def gradient(P, F, T, mu, sd):
    """Batch MSE and its gradient from the library's backward kernel
    (no parameter update): (loss, grad) as device tensors."""
    dev = F.device
    grad = torch.zeros(NPARAM, dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    ctx = engine.context(dev)
    rc = lib.mhip_mlp_train_step(ctx, engine.ptr(F), engine.ptr(T), F.shape[0], engine.ptr(mu),
                                 engine.ptr(sd), engine.ptr(P), None, None, engine.ptr(grad),
                                 engine.ptr(loss), 0, 0.0)
    raise_for(rc, ctx)
    return loss, grad
