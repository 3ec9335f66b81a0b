"""``Meshoid`` -- drop-in for the reference class's neighbour-weighted operator
path (src/meshoid/Meshoid.py:12-331), computed by libmeshoid_hip.so.

Same constructor (``Meshoid(pos, m=None, kernel_radius=None, des_ngb=None,
boxsize=None, verbose=False, particle_mask=None, n_jobs=-1)``; ``hsml=`` is an
alias of ``kernel_radius``), same methods and attributes.  Attributes that the
reference stores as NumPy arrays (``ngb``, ``ngbdist``, ``kernel_radius``,
``density``, ``vol``, ``dweights``, ``d2weights``, ``dweights_3rdorder``) are
produced from device results on first access; ``*_device`` accessors return
the device tensors.  Known quirks of the reference are kept:

* ``TreeUpdate`` multiplies ``self.m[self.particle_mask]`` although ``self.m``
  is already masked (Meshoid.py:92 vs :169);
* ``Curl``/``Div`` pass ``v[self.particle_mask]`` to ``D``, which indexes it
  with global neighbour indices (Meshoid.py:310,330): IndexError when out of
  range, as ``np.take`` raises;
* ``Curl`` returns (dvy/dz - dvz/dy, dvx/dz - dvz/dx, dvx/dy - dvy/dx)
  (Meshoid.py:311-315); ``Curl(v, correct_sign=True)`` gives the textbook curl;
* the ``weighted`` flag of ``D``/``D2`` only matters for the first call of
  each order, because the reference caches the weights (Meshoid.py:261-267);
* ``Meshoid(pos, m, kernel_radius=h).D(f)`` re-runs the neighbour search and
  replaces ``h`` by the adaptive smoothing length (Meshoid.py:255-256 -> 165).
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine
from ._lib import ST_SINGULAR
from .functions import as_device, gradient_weights as _gradient_weights

_UNSET = object()


class Meshoid:
    """Meshoid object that stores particle data and kernel-weighting
    quantities, and implements derivative operators on the GPU."""

    def __init__(self, pos, m=None, kernel_radius=None, des_ngb=None, boxsize=None, verbose=False,
                 particle_mask=None, n_jobs=-1, *, hsml=None):
        if kernel_radius is None and hsml is not None:
            kernel_radius = hsml
        self.tree = None
        if len(pos.shape) == 1:
            pos = pos[:, None]
        self.verbose = verbose
        self.N, self.dim = pos.shape
        if particle_mask is None:
            self.particle_mask = np.arange(self.N)
        else:
            pm = particle_mask.cpu().numpy() if isinstance(particle_mask, torch.Tensor) else particle_mask
            if np.array(pm).dtype == np.dtype("bool"):
                self.particle_mask = np.arange(len(pm))[pm]
            else:
                self.particle_mask = np.asarray(pm)
        self.Nmask = len(self.particle_mask)
        if des_ngb is None:
            des_ngb = {1: 4, 2: 20, 3: 32}[self.dim]
        self.des_ngb = des_ngb
        self.n_jobs = n_jobs
        self.volnorm = {1: 2.0, 2: np.pi, 3: 4 * np.pi / 3}[self.dim]
        self.boxsize = boxsize
        self.pos = pos
        self._pos_d = as_device(pos)
        self._center = self._L = None
        if m is None:
            m = np.repeat(1.0, self.N)
        m = m.cpu().numpy() if isinstance(m, torch.Tensor) else np.asarray(m)
        self.m = m[self.particle_mask]
        self._identity_mask = self.Nmask == self.N and np.array_equal(
            self.particle_mask, np.arange(self.N))
        self._nl = None
        self._nl_full = None  # the unmasked list of the current tree, kept for masked updates
        self._dev = {}
        self._np = {}
        self._dw = {}
        self._weighted = {}
        self.sliceweights = None
        if kernel_radius is None:
            self.TreeUpdate()
        else:
            h = kernel_radius.cpu().numpy() if isinstance(kernel_radius, torch.Tensor) else kernel_radius
            self._np["kernel_radius"] = h
            self._np["vol"] = self.volnorm * h**self.dim / self.des_ngb
            self._np["density"] = self.m / self._np["vol"]

    # ------------------------------------------------------------------ state
    def _cached(self, name):
        if name in self._np:
            return self._np[name]
        if name in self._dev:
            a = self._dev[name].cpu().numpy().astype(np.float64)
            self._np[name] = a
            return a
        return None

    kernel_radius = property(lambda s: s._cached("kernel_radius"),
                             lambda s, v: s._set("kernel_radius", v))
    density = property(lambda s: s._cached("density"), lambda s, v: s._set("density", v))
    vol = property(lambda s: s._cached("vol"), lambda s, v: s._set("vol", v))

    def _set(self, name, v):
        self._dev.pop(name, None)
        self._np[name] = v

    @property
    def ngb(self):
        if "ngb" in self._np:
            return self._np["ngb"]
        if self._nl is None:
            return None
        self._export()
        return self._np["ngb"]

    @ngb.setter
    def ngb(self, v):
        self._np["ngb"] = v
        if v is None:
            self._nl = None

    @property
    def ngbdist(self):
        if "ngbdist" in self._np:
            return self._np["ngbdist"]
        if self._nl is None:
            raise AttributeError("'Meshoid' object has no attribute 'ngbdist'")
        self._export()
        return self._np["ngbdist"]

    @ngbdist.setter
    def ngbdist(self, v):
        self._np["ngbdist"] = v

    def _export(self):
        ngb, dist = self._nl.export(f64=True)
        self._dev["ngb"], self._dev["ngbdist"] = ngb, dist
        self._np["ngb"] = ngb.cpu().numpy().astype(np.int64)
        self._np["ngbdist"] = dist.cpu().numpy().astype(np.float64)

    @property
    def center(self):
        if self._center is None:
            self._frame()
        return self._center

    @center.setter
    def center(self, v):
        self._center = v

    @property
    def L(self):
        if self._L is None:
            self._frame()
        return self._L

    @L.setter
    def L(self, v):
        self._L = v

    def _frame(self):
        """center/L of Meshoid.py:81-87 (computed on first use)."""
        if self.boxsize is None:
            p = np.asarray(self.pos.cpu() if isinstance(self.pos, torch.Tensor) else self.pos)
            self._center = np.median(p, axis=0)
            self._L = 2 * np.percentile(np.sum((p - self._center) ** 2, axis=1), 90) ** 0.5
        else:
            self._center = np.ones(3) * self.boxsize / 2
            self._L = self.boxsize

    def device_arrays(self):
        """Device tensors of the latest TreeUpdate: h, rho, vol, status."""
        return dict(self._dev)

    # ------------------------------------------------------------ hot path
    def reset_dweights(self):
        self._dw.clear()
        self._weighted.clear()

    def BuildTree(self):
        if self.verbose:
            print("Building tree...")
        self.tree = engine.SpatialIndex().build(self._pos_d, self.boxsize)
        self._nl_full = None

    def TreeUpdate(self):
        """Neighbour lists, smoothing lengths and densities (Meshoid.py:147-170)."""
        if self.tree is None:
            self.BuildTree()
        if self.verbose:
            print("Finding neighbors...")
        dev = self._pos_d.device
        m_rows = np.asarray(self.m[self.particle_mask], dtype=np.float64)  # Meshoid.py:169
        m_d = torch.from_numpy(np.ascontiguousarray(m_rows)).to(dev)
        qidx = None
        if not self._identity_mask:
            pm = np.asarray(self.particle_mask, dtype=np.int64)
            if pm.size and (pm.min() < -self.N or pm.max() >= self.N):
                raise IndexError("particle_mask index out of range")
            qidx = torch.from_numpy(np.where(pm < 0, pm + self.N, pm)).to(dev)
        full = self._nl_full
        if qidx is not None and full is not None and full.k == self.des_ngb and full.index is self.tree:
            # the rows of the unmasked list are the masked search's result
            if self._nl is full:
                self._nl = None
            out = engine.nlist_subset(full, qidx, m_rows=m_d, nlist=self._nl)
        else:
            if qidx is not None and self._nl is full:
                self._nl = None  # keep the unmasked list intact
            # ngbdist is exported in float64 from the positions: the list keeps no distances
            out = engine.knn_hsml(self.tree, self.des_ngb, qidx=qidx, m_rows=m_d, keep_dist=False,
                                  nlist=self._nl)
            if qidx is None:
                self._nl_full = out["nlist"]
        if self.verbose:
            print("Neighbours found!")
            print("Iterating for smoothing lengths...")
            print("Smoothing lengths found!")
        self._nl = out["nlist"]
        for key in ("ngb", "ngbdist", "kernel_radius", "density", "vol"):
            self._np.pop(key, None)
        # float64 h of the final neighbours (HsmlIter's root to ~1e-13), and
        # density and volume from it in float64, as Meshoid.py:165-170
        h = self._nl.h64()
        rho = self.des_ngb * m_d / (self.volnorm * h**self.dim)
        self._dev = {"kernel_radius": h, "density": rho, "vol": m_d / rho, "status": out["status"]}

    def get_kernel_weights(self):
        """W(r_ij / h_i) normalised over each row's neighbours
        (Meshoid.py:172-177): (Nmask, k) float64, columns as ``ngbdist``."""
        if self.ngbdist is None or self.kernel_radius is None:
            self.TreeUpdate()
        if self._nl is not None and "ngbdist" not in self._dev:
            self._export()
        dist = self._dev.get("ngbdist")
        h = self._dev.get("kernel_radius")
        dev = self._pos_d.device
        if dist is None:  # lists or radii assigned by the caller
            dist = torch.from_numpy(np.ascontiguousarray(self.ngbdist, dtype=np.float64)).to(dev)
        if h is None or "kernel_radius" in self._np:
            h = torch.from_numpy(np.ascontiguousarray(self.kernel_radius, dtype=np.float64)).to(dev)
        return engine.kernel_weights_dev(dist.contiguous(), h.contiguous()).cpu().numpy()

    def _need_device_list(self, what):
        if self._nl is None:
            raise NotImplementedError(f"{what} runs on the GPU neighbour list of TreeUpdate(); "
                                      "lists assigned by the caller are not supported")

    def KernelAverage(self, f):
        """Kernel-weighted average of a scalar field over each row's
        neighbours (Meshoid.py:369-381): (Nmask,) float64."""
        if self.ngbdist is None or self.kernel_radius is None:  # via get_kernel_weights
            self.TreeUpdate()
        self._need_device_list("KernelAverage")
        is_t = isinstance(f, torch.Tensor)
        if len(f.shape) != 1:
            # einsum("ij,ij->i") of the reference rejects extra axes
            raise ValueError("operands could not be broadcast: KernelAverage takes f of shape (N,)")
        fd, _, _ = self._field(f)
        out = engine.kernel_average(self._nl, fd.reshape(-1).contiguous())
        return out if is_t else out.cpu().numpy()

    def KernelVariance(self, f):
        """Standard deviation of f over each row's neighbours
        (Meshoid.py:352-367): (Nmask, ...)."""
        if self.ngb is None:
            self.TreeUpdate()
        self._need_device_list("KernelVariance")
        fd, tail, is_t = self._field(f)
        out = engine.kernel_variance(self._nl, fd).reshape((self.Nmask,) + tuple(tail))
        if is_t:
            return out.to(f.dtype) if f.dtype in (torch.float32, torch.float64) else out
        dt = np.asarray(f).dtype
        return out.cpu().numpy().astype(dt if dt in (np.float32, np.float64) else np.float64)

    def Integrate(self, f):
        """Volume integral sum_i vol_i f_i over the masked particles
        (Meshoid.py:333-350): shape f.shape[1:], float64."""
        if self.kernel_radius is None:
            self.TreeUpdate()
        elif self.vol is None:
            self.vol = self.volnorm * self.kernel_radius**self.dim
        dev = self._pos_d.device
        vol = self._dev.get("vol") if "vol" not in self._np else None
        if vol is None:  # set by the caller or from kernel_radius: float64, as the reference's
            vol = torch.from_numpy(np.ascontiguousarray(self.vol, dtype=np.float64)).to(dev)
        is_t = isinstance(f, torch.Tensor)
        tail = tuple(f.shape[1:])
        fd = as_device(f)
        if not self._identity_mask:
            fd = fd[torch.as_tensor(self.particle_mask, device=dev)]
        out = engine.integrate(vol.contiguous(), fd.reshape(fd.shape[0], -1).contiguous()).reshape(tail)
        return out if is_t else out.cpu().numpy()

    def Volume(self):
        return self.vol

    def NearestNeighbors(self):
        if self.ngb is None:
            self.TreeUpdate()
        return self.ngb

    def NeighborDistance(self):
        if self.ngbdist is None:
            self.TreeUpdate()
        return self.ngbdist

    def SmoothingLength(self):
        if self.kernel_radius is None:
            self.TreeUpdate()
        return self.kernel_radius

    def Density(self):
        if self.density is None:
            self.TreeUpdate()
        return self.density

    # ------------------------------------------------------ derivatives
    def ComputeDWeights(self, order=1, weighted=True):
        """Materialise the derivative weights (Meshoid.py:110-140)."""
        if self.verbose:
            print(f"Computing weights for derivatives of order {order}...")
        self._weighted[order] = bool(weighted)
        ngb = self.ngb
        w = _gradient_weights(self._pos_d, as_device(ngb, (torch.int32,), torch.int32),
                              as_device(self.kernel_radius, (torch.float64,)),
                              as_device(self.particle_mask, (torch.int64,), torch.int64),
                              boxsize=self.boxsize, weighted=weighted, order=order)
        w = w.cpu().numpy()
        if order == 1:
            self._dw[1] = w
        elif order == 2:
            self._dw[2] = (w[:, self.dim:, :], w[:, : self.dim, :])

    dweights = property(lambda s: s._dw.get(1), lambda s, v: s._dw.__setitem__(1, v))
    d2weights = property(lambda s: s._dw[2][0] if 2 in s._dw else None)
    dweights_3rdorder = property(lambda s: s._dw[2][1] if 2 in s._dw else None)

    def _field(self, f):
        """(nf, K) device tensor + tail shape; checks the np.take bounds."""
        is_t = isinstance(f, torch.Tensor)
        shape = tuple(f.shape)
        if len(shape) == 0:
            raise ValueError("f must have a leading particle axis")
        fd = as_device(f)
        fd = fd.reshape(shape[0], -1).contiguous()
        nf = shape[0]
        if nf < self.N:
            # the reference indexes f with global neighbour indices (np.take)
            if self.ngb is not None and self.ngb.size and int(self.ngb.max()) >= nf:
                raise IndexError(f"index {int(self.ngb.max())} is out of bounds for axis 0 with size {nf}")
            if self.particle_mask.size and int(np.max(self.particle_mask)) >= nf:
                raise IndexError(f"index {int(np.max(self.particle_mask))} is out of bounds for axis 0 with size {nf}")
        return fd, shape[1:], is_t

    def _apply_dev(self, fd, order, weighted, second):
        """(Nmask, K, nd) float32 device derivatives of the (nf, K) field."""
        w = self._weighted.setdefault(order, bool(weighted))
        if self.verbose and order not in self._dw:
            print(f"Computing weights for derivatives of order {order}...")
        d1, d2, st = engine.apply_D(self._nl, fd, order=order, weighted=w, want_d1=not second,
                                    want_d2=second)
        if bool(((st & ST_SINGULAR) != 0).any()):
            raise np.linalg.LinAlgError("Matrix is singular to machine precision.")
        return d2 if second else d1

    def _apply(self, f, order, weighted, second):
        if self._nl is None:
            self.TreeUpdate()
        fd, tail, is_t = self._field(f)
        out = self._apply_dev(fd, order, weighted, second)
        out = out.reshape((self.Nmask,) + tuple(tail) + (out.shape[-1],))
        if is_t:
            return out
        return out.cpu().numpy().astype(np.float64)

    def D(self, f, order=1, weighted=True):
        """Kernel-weighted least-squares gradient of f (Meshoid.py:237-268):
        (Nmask, ..., dim)."""
        return self._apply(f, order, weighted, second=False)

    def D2(self, f, weighted=True):
        """Least-squares second derivatives (Meshoid.py:270-295):
        (Nmask, ..., nd2), 3D order [xx, yy, zz, xy, yz, zx]."""
        return self._apply(f, 2, weighted, second=True)

    def Curl(self, v, correct_sign=False):
        """Curl of a vector field (Meshoid.py:297-315).  The reference's
        component signs are kept unless correct_sign=True."""
        dv = self.D(v[self.particle_mask])
        c = [dv[:, 1, 2] - dv[:, 2, 1], dv[:, 0, 2] - dv[:, 2, 0], dv[:, 0, 1] - dv[:, 1, 0]]
        if correct_sign:
            c = [-c[0], c[1], -c[2]]
        if isinstance(dv, torch.Tensor):
            return torch.stack(c, dim=1)
        return np.c_[c[0], c[1], c[2]]

    def Div(self, v):
        """Divergence of a vector field (Meshoid.py:317-331)."""
        dv = self.D(v[self.particle_mask])
        return dv[:, 0, 0] + dv[:, 1, 1] + dv[:, 2, 2]

    # ---------------------------------------------------- reconstruction
    def _set_mask(self, idx):
        self.particle_mask = idx
        self.Nmask = len(idx)
        self._identity_mask = self.Nmask == self.N and np.array_equal(idx, np.arange(self.N))

    def Reconstruct(self, f, target_points, order: int = 1):
        """Values of f at arbitrary points from each point's nearest particle
        (Meshoid.py:383-440): order 0 its value, 1 plus dx . D(f), 2 (3D) plus
        the second-order terms of D2(f).  As in the reference, order >= 1 sets
        ``particle_mask`` to the sorted distinct nearest particles and re-runs
        ``TreeUpdate`` on them when that changes it.  Returns (N_target, ...)
        in f's dtype."""
        if self.tree is None:
            self.BuildTree()
        dev = self._pos_d.device
        tp = target_points
        tp_d = tp.to(dev, torch.float64) if isinstance(tp, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(tp, dtype=np.float64)).to(dev)
        single = tp_d.dim() == 1 and self.dim > 1
        tp_d = tp_d.reshape(-1, self.dim if single or tp_d.dim() == 1 else tp_d.shape[-1]).contiguous()
        nbr = engine.nearest(self.tree, tp_d)
        is_t = isinstance(f, torch.Tensor)
        tail = tuple(f.shape[1:])
        # f keeps its dtype (order 0 returns np.take(f, ...) of any dtype)
        fd = (f.to(dev) if is_t else torch.from_numpy(np.ascontiguousarray(f)).to(dev)).contiguous()
        nf = fd.shape[0]
        fd = fd.reshape(nf, -1).contiguous()
        floating = fd.dtype in (torch.float32, torch.float64)
        if nbr.numel() and nf < self.N and int(nbr.max()) >= nf:
            raise IndexError(f"index {int(nbr.max())} is out of bounds for axis 0 with size {nf}")

        def done(out):
            out = out.reshape((tp_d.shape[0],) + tail)
            if single:
                out = out[0]
            return out if is_t else out.cpu().numpy()

        if order == 0:
            if not floating:  # np.take of any dtype
                if nbr.numel() and int(nbr.min()) < 0:
                    raise ValueError("target_points must be finite")
                return done(fd[nbr])
            inv = torch.zeros_like(nbr)
            return done(engine.reconstruct(tp_d, self._pos_d, self.boxsize, nbr, inv, fd, order=0))
        if order not in (1, 2):
            raise ValueError("order must be 0, 1 or 2")
        uniq, inv = engine.unique_inverse(nbr, self.N)
        u = uniq.cpu().numpy()
        if not np.array_equal(self.particle_mask, u):
            self._set_mask(u)
            self.TreeUpdate()  # neighbour lists of just the target neighbours
            self.reset_dweights()
        if not floating:
            # the reference's in-place f_target += float64 fails the same way
            raise TypeError(f"Cannot cast ufunc 'add' output from dtype('float64') to "
                            f"dtype('{np.dtype(str(fd.dtype).split('.')[-1])}') with casting rule 'same_kind'")
        if order == 2 and self.dim != 3:
            raise ValueError("order-2 reconstruction is defined for 3D particles only")
        if self._nl is None:
            self.TreeUpdate()
        d1 = self._apply_dev(fd, order, True, second=False).contiguous()
        d2 = self._apply_dev(fd, 2, True, second=True).contiguous() if order == 2 else None
        return done(engine.reconstruct(tp_d, self._pos_d, self.boxsize, nbr, inv, fd, d1, d2, order))

    def Slice(self, f, size=None, plane="z", center=None, res: int = 128, order: int = 1,
              return_grid: bool = False):
        """f reconstructed on a res x res grid in a plane through ``center``
        (Meshoid.py:442-502): the grid spans [-size/2, size/2] (end points
        included) along the two in-plane axes, normal ``plane`` ('x', 'y',
        otherwise z)."""
        if center is None:
            center = self.center
        if size is None:
            size = self.L
        if self.tree is None:
            self.TreeUpdate()
        ax = np.linspace(-size / 2, size / 2, res)
        x, y = np.meshgrid(ax, ax, indexing="ij")
        u, v, w = x.ravel(), y.ravel(), np.zeros(res * res)
        cols = {"x": (w, u, v), "y": (u, w, v)}.get(plane, (u, v, w))
        grid = np.column_stack(cols) + center
        f_grid = self.Reconstruct(f, grid, order)
        shape = (res, res) + tuple(f.shape[1:])
        f_grid = f_grid.reshape(shape)
        if return_grid:
            return x, y, f_grid
        return f_grid

    def InterpToGrid(self, f, size=None, center=None, res: int = 128, order: int = 1,
                     return_grid: bool = False):
        """f reconstructed at the cell centres of a res^3 grid of side
        ``size`` around ``center`` (Meshoid.py:504-555)."""
        if center is None:
            center = self.center
        if size is None:
            size = self.L
        edges = np.linspace(-size / 2, size / 2, res + 1)
        mid = (edges[1:] + edges[:-1]) / 2
        X, Y, Z = np.meshgrid(mid, mid, mid, indexing="ij")
        grid = np.column_stack((X.ravel(), Y.ravel(), Z.ravel())) + center
        f_grid = self.Reconstruct(f, grid, order).reshape((res, res, res) + tuple(f.shape[1:]))
        if return_grid:
            return X, Y, Z, f_grid
        return f_grid

    # ---------------------------------------------------------- projection
    def SurfaceDensity(self, f=None, size=None, center=None, res=128, smooth_fac=1.0, conservative=False):
        """Surface density of f (default: the particle masses) on a res x res
        grid of sightlines along z (Meshoid.py:595-643): GridSurfaceDensity of
        f at the particle positions with smoothing lengths
        clip(smooth_fac * h, 2 * size / res, 1e100), float64 (res, res)."""
        if f is None:
            f = self.m
        if center is None:
            center = self.center
        if size is None:
            size = self.L
        if self.dim < 2:
            raise IndexError("SurfaceDensity projects (x, y): positions need at least 2 dimensions")
        if self.kernel_radius is None and "kernel_radius" not in self._dev:
            self.TreeUpdate()
        dev = self._pos_d.device
        h = self._dev.get("kernel_radius") if "kernel_radius" not in self._np else None
        if h is None:
            h = torch.from_numpy(np.ascontiguousarray(self.kernel_radius, dtype=np.float64)).to(dev)
        is_t = isinstance(f, torch.Tensor)
        fd = as_device(f).reshape(-1).contiguous()
        grid = engine.surface_density(fd, self._pos_d, h.contiguous(), np.asarray(center, dtype=np.float64),
                                      float(size), int(res), float(smooth_fac), bool(conservative))
        return grid if is_t else grid.cpu().numpy()

    def DepositToGrid(self, f, weights=None, size=None, center=None, res=128):
        """Density of the conserved quantity f on a res^3 grid
        (Meshoid.py:557-593): GridDensity with h = clip(h, size/(res-1), 1e100),
        indices wrapped periodically when the meshoid has a box.  ``weights``
        is accepted and unused, as in the reference.  Without a box the
        reference's call fails in its compiled kernel (box_size None); here the
        grid indices are clipped, as GridDensity does for box_size < 0."""
        if center is None:
            center = self.center
        if size is None:
            size = self.L
        if self.dim != 3:
            raise ValueError("DepositToGrid needs 3D positions")
        if self.kernel_radius is None and "kernel_radius" not in self._dev:
            self.TreeUpdate()
        dev = self._pos_d.device
        h = self._dev.get("kernel_radius") if "kernel_radius" not in self._np else None
        if h is None:
            h = torch.from_numpy(np.ascontiguousarray(self.kernel_radius, dtype=np.float64)).to(dev)
        is_t = isinstance(f, torch.Tensor)
        fd = as_device(f).reshape(-1).contiguous()
        g = engine.grid_density(fd, self._pos_d, h.contiguous(), np.asarray(center, dtype=np.float64), float(size),
                                int(res), self.boxsize is not None, float(size) / (res - 1))
        return g if is_t else g.cpu().numpy()
