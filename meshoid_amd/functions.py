"""Module-level functions of the reference's star-exported API, GPU-backed.

* :func:`HsmlIter`          -- src/meshoid/kernel_density.py:14-57
* :func:`Kernel`            -- src/meshoid/kernel_density.py:82-96
* :func:`gradient_weights`  -- src/meshoid/derivatives.py:117-138
* :func:`nearest_image`, :func:`nearest_image_v` -- src/meshoid/derivatives.py:8-23

NumPy inputs give NumPy outputs (float64, like the reference); device
tensors give device tensors.
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine
from ._lib import ST_HSML_NOCONV, ST_SINGULAR


def _device():
    return torch.device("cuda", torch.cuda.current_device())


def as_device(x, dtypes=(torch.float32, torch.float64), default=torch.float64):
    """Device tensor view/copy of x keeping float32/float64 (others -> default)."""
    if isinstance(x, torch.Tensor):
        t = x if x.is_cuda else x.to(_device())
    else:
        a = np.asarray(x)
        t = torch.from_numpy(np.ascontiguousarray(a)).to(_device())
    if t.dtype not in dtypes:
        t = t.to(default)
    return t.contiguous()


def _out(t, like, dtype=np.float64):
    if isinstance(like, torch.Tensor):
        return t
    return t.cpu().numpy().astype(dtype, copy=False)


def HsmlIter(neighbor_dists, dim=3, error_norm=1e-6):
    """Smoothing lengths from (N, Nngb) ascending neighbour distances: the root
    of norm_d * sum_j W(r_ij/h) = Nngb (kernel_density.py:14-57)."""
    if dim not in (1, 2, 3):
        raise KeyError(dim)
    d = as_device(neighbor_dists)
    h, status = engine.hsml_iter_dev(d, dim=dim, error_norm=error_norm)
    if bool(((status & ST_HSML_NOCONV) != 0).any()):
        import warnings

        warnings.warn("HsmlIter: some rows did not reach error_norm within the iteration cap")
    return _out(h, neighbor_dists)


def Kernel(q):
    """Un-normalised cubic spline W(q) (kernel_density.py:82-96)."""
    scalar = np.isscalar(q)
    t = as_device(np.atleast_1d(q) if scalar else q)
    w = engine.kernel_dev(t)
    if scalar:
        return float(w.cpu()[0])
    if isinstance(q, torch.Tensor):
        return w
    return w.cpu().numpy()


def nearest_image(dx_coord, boxsize):
    """Separation for the nearest periodic image (derivatives.py:8-14)."""
    if abs(dx_coord) > boxsize / 2:
        return -np.copysign(boxsize - abs(dx_coord), dx_coord)
    return dx_coord


def nearest_image_v(dx_coord, boxsize):
    """Elementwise nearest_image (derivatives.py:17-23)."""
    if isinstance(dx_coord, torch.Tensor):
        a = dx_coord.abs()
        return torch.where(a > boxsize / 2, -torch.copysign(boxsize - a, dx_coord), dx_coord)
    dx = np.asarray(dx_coord)
    a = np.abs(dx)
    return np.where(a > boxsize / 2, -np.copysign(boxsize - a, dx), dx)


def gradient_weights(pos, ngb, kernel_radius, indices, boxsize=None, weighted=True, order=1):
    """(N, nd, Nngb) float64 least-squares derivative weights
    (derivatives.py:117-138).  Raises LinAlgError on an exactly singular
    moment matrix, as np.linalg.inv does."""
    p = as_device(pos)
    if p.dim() == 1:
        p = p[:, None]
    if p.shape[1] not in (1, 2, 3):
        raise KeyError(p.shape[1])
    g = as_device(ngb, dtypes=(torch.int32,), default=torch.int32)
    h = as_device(kernel_radius, dtypes=(torch.float64,))
    idx = as_device(indices, dtypes=(torch.int64,), default=torch.int64)
    G, status = engine.gradient_weights_dev(p, g, h, idx, boxsize, weighted, order)
    if bool(((status & ST_SINGULAR) != 0).any()):
        raise np.linalg.LinAlgError("Matrix is singular to machine precision.")
    return _out(G, pos)


def _dev_like(a, dtypes=(torch.float32, torch.float64)):
    return as_device(a, dtypes).contiguous()


def GridSurfaceDensity(f, x, h, center, size, res=100, box_size=-1, parallel=True, conservative=False):
    """Surface density of the conserved quantity f at positions x (N, 2|3)
    with smoothing lengths h on a res x res grid of sightlines along z
    (grid_deposition.py:123-295): float64 (res, res), points at
    center - size/2 + g size/(res-1).  box_size and parallel are accepted and,
    as in the reference's cores, unused."""
    pos = _dev_like(x)
    if pos.dim() != 2 or pos.shape[1] < 2:
        raise IndexError("GridSurfaceDensity needs (N, 2) or (N, 3) positions")
    g = engine.surface_density(_dev_like(f).reshape(-1), pos, _dev_like(h).reshape(-1),
                               np.asarray(center, dtype=np.float64), float(size), int(res), 1.0,
                               bool(conservative), hmin=0.0)
    return _out(g, x)


def GridDensity(f, x, h, center, size, res=100, box_size=-1.0):
    """Density of the conserved quantity f on a res^3 grid
    (grid_deposition.py:534-603): float64 (res, res, res), points at
    center - size/2 + dx/2 + g dx, dx = size/(res-1); box_size >= 0 wraps the
    grid indices periodically, as the reference does."""
    periodic = box_size is not None and box_size >= 0
    g = engine.grid_density(_dev_like(f).reshape(-1), _dev_like(x), _dev_like(h).reshape(-1),
                            np.asarray(center, dtype=np.float64), float(size), int(res), periodic, 0.0)
    return _out(g, x)
