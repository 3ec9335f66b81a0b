"""Radiative transfer through the particles on the GPU: ``radtransfer``
(src/meshoid/radiation/radtransfer.py:11-177), computed by
libmeshoid_hip.so (``mhip_radtransfer``)."""
from __future__ import annotations

import numpy as np
import torch

from . import engine
from ._lib import lib, raise_for
from .functions import as_device


def radtransfer(j, m, kappa, x, h, gridres, L, center=0, i0=0):
    """Emission and absorption along z on a gridres x gridres grid of
    sightlines centred on ``center`` with side L (points at
    center + (-L/2 + g L/(gridres-1))): returns the intensity
    (gridres, gridres, num_freqs), float64.  j (N, num_freqs) emissivity per
    unit mass, m (N,), kappa (N, num_freqs) opacity ((N,) is taken for every
    band), x (N, 3), h (N,), i0 the background: a scalar, (num_freqs,) or
    (gridres, gridres, num_freqs).  The particles act farthest first (largest
    z first), as the reference's single-chunk sweep orders them; the
    reference's multi-threaded driver orders its chunks the other way and
    adds i0 once per chunk, so it matches this with one thread."""
    dev = torch.device("cuda", torch.cuda.current_device())
    f64 = (torch.float64,)
    xd = as_device(x, f64, torch.float64).to(torch.float64).reshape(-1, 3).contiguous()
    n = xd.shape[0]
    jd = as_device(j, f64, torch.float64).to(torch.float64)
    if jd.dim() == 1:
        jd = jd[:, None]
    nf = jd.shape[1]
    kd = as_device(kappa, f64, torch.float64).to(torch.float64)
    if kd.dim() == 1:
        kd = kd[:, None].expand(-1, nf)
    jd, kd = jd.contiguous(), kd.contiguous()
    md = as_device(m, f64, torch.float64).to(torch.float64).reshape(-1).contiguous()
    hd = as_device(h, f64, torch.float64).to(torch.float64).reshape(-1).contiguous()
    if jd.shape[0] != n or kd.shape != (n, nf) or md.numel() != n or hd.numel() != n:
        raise ValueError("j, m, kappa, x and h must have one row per particle")
    c = np.broadcast_to(np.asarray(center, dtype=np.float64), (3,))
    res = int(gridres)
    i0a = np.asarray(i0, dtype=np.float64)
    i0d = None
    if np.any(i0a != 0):
        # a C-ordered, writable copy: the library reads the buffer as (res, res, nf)
        i0d = torch.from_numpy(np.array(np.broadcast_to(i0a, (res, res, nf)), dtype=np.float64, order="C"))
        i0d = i0d.to(dev).contiguous()
    out = torch.empty((res, res, nf), dtype=torch.float64, device=dev)
    ctx = engine.context(dev)
    raise_for(lib.mhip_radtransfer(ctx, engine.ptr(jd), engine.ptr(md), engine.ptr(kd), engine.ptr(xd),
                                   engine.ptr(hd), n, nf, res, float(L), float(c[0]), float(c[1]),
                                   engine.ptr(i0d), engine.ptr(out)), ctx)
    return out.cpu().numpy()


__all__ = ["radtransfer"]
