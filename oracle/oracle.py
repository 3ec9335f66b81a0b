"""FP64 CPU oracle for meshoid's neighbour-weighted operator path.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by ``__graft_entry__.smoke()``
(as the checker) and by the ``cpu_baseline`` leg of ``bench.py``.  The product
package ``meshoid_amd`` never imports it.

Each function restates the reference code it cites (paths relative to the
meshoid source tree):

* :func:`knn`                -- ``cKDTree(pos, boxsize).query(x, k)``, the call at
  ``src/meshoid/Meshoid.py:145,157`` (exact kNN, self included, ascending).
* :func:`hsml_iter`          -- ``HsmlIter``, ``src/meshoid/kernel_density.py:14-57``.
* :func:`kernel`             -- ``Kernel``, ``src/meshoid/kernel_density.py:82-96``.
* :func:`gradient_weights`   -- ``gradient_weights``, ``src/meshoid/derivatives.py:117-138``.
* :func:`apply_weights`      -- the einsum of ``Meshoid.D``/``D2``,
  ``src/meshoid/Meshoid.py:258,268,291,295``.
* :class:`OracleMeshoid`     -- the ``Meshoid`` state machine for the hot-path
  methods (``Meshoid.py:15-105,110-170,237-331``).
* :func:`kernel_weights`, :func:`kernel_average`, :func:`kernel_variance`,
  :func:`integrate` -- ``get_kernel_weights``, ``KernelAverage``,
  ``KernelVariance`` and ``Integrate``, ``src/meshoid/Meshoid.py:172-177,333-381``.
* :func:`grid_surface_density`, :func:`kernel2d` -- ``GridSurfaceDensity`` and its
  serial cores, ``src/meshoid/grid_deposition.py:123-295``, ``kernel2d``
  ``src/meshoid/kernel_density.py:60-68``; :meth:`OracleMeshoid.SurfaceDensity`
  (``Meshoid.py:595-643``).
* :func:`grid_density`, :func:`kernel3d` -- ``GridDensity``
  ``src/meshoid/grid_deposition.py:534-603``, ``kernel3d``
  ``src/meshoid/kernel_density.py:71-79``; :meth:`OracleMeshoid.DepositToGrid`
  (``Meshoid.py:557-593``).
* :func:`radtransfer` -- ``radtransfer_singlethread``,
  ``src/meshoid/radiation/radtransfer.py:79-177`` (farthest first).
* :meth:`OracleMeshoid.Reconstruct`, ``Slice``, ``InterpToGrid`` --
  ``src/meshoid/Meshoid.py:383-555`` (the nearest particle from :func:`knn`
  with k = 1, ties to the lowest index).

The numerical core is ``meshoid_oracle.c`` (strict FP64, no FMA contraction),
built into ``liboracle.so`` next to this file.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so in place (gcc, see Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_knn.argtypes = [_f64p, ctypes.c_int64, ctypes.c_int, ctypes.c_double, _f64p,
                              ctypes.c_int64, ctypes.c_int, _i64p, _f64p]
        L.orc_knn_brute.argtypes = L.orc_knn.argtypes
        L.orc_hsml_iter.argtypes = [_f64p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int, _f64p, _i32p]
        L.orc_kernel.argtypes = [ctypes.c_double]
        L.orc_kernel.restype = ctypes.c_double
        L.orc_gradient_weights.argtypes = [_f64p, ctypes.c_int64, ctypes.c_int, _i64p,
                                           ctypes.c_int64, ctypes.c_int, _f64p, _i64p,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, _f64p,
                                           _i32p]
        L.orc_apply_weights.argtypes = [_f64p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _i64p, _i64p, _f64p,
                                        ctypes.c_int, _f64p]
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def num_threads() -> int:
    return int(lib().orc_num_threads())


def _pos2d(pos):
    pos = np.asarray(pos, dtype=np.float64)
    if pos.ndim == 1:
        pos = pos[:, None]
    return np.ascontiguousarray(pos)


def knn(pos, k, boxsize=None, queries=None, brute=False):
    """Exact kNN (f64).  Returns (dist (nq,k) f64, ngb (nq,k) int64)."""
    pos = _pos2d(pos)
    n, dim = pos.shape
    q = pos if queries is None else _pos2d(queries)
    nq = q.shape[0]
    ngb = np.empty((nq, k), dtype=np.int64)
    dist = np.empty((nq, k), dtype=np.float64)
    fn = lib().orc_knn_brute if brute else lib().orc_knn
    rc = fn(pos, n, dim, float(boxsize) if boxsize else -1.0, q, nq, int(k), ngb, dist)
    if rc != 0:
        raise ValueError("oracle knn failed")
    return dist, ngb


def hsml_iter(neighbor_dists, dim=3, error_norm=1e-6, max_iter=1000, return_iters=False):
    d = np.ascontiguousarray(neighbor_dists, dtype=np.float64)
    n, k = d.shape
    h = np.empty(n, dtype=np.float64)
    it = np.empty(n, dtype=np.int32)
    lib().orc_hsml_iter(d, n, k, dim, float(error_norm), int(max_iter), h, it)
    return (h, it) if return_iters else h


def kernel(q):
    q = np.asarray(q, dtype=np.float64)
    return np.where(q <= 0.5, 1 - 6 * q**2 + 6 * q**3, np.where(q <= 1.0, 2 * (1 - q) ** 3, 0.0))


def kernel_weights(ngbdist, kernel_radius):
    """Kernel(ngbdist / h) normalised per row (Meshoid.py:172-177)."""
    K = kernel(np.asarray(ngbdist, np.float64) / np.asarray(kernel_radius, np.float64)[:, None])
    return K / np.sum(K, axis=1)[:, None]


def kernel_average(ngbdist, kernel_radius, ngb, f):
    """einsum("ij,ij->i", weights, f[ngb]) (Meshoid.py:369-381), scalar f."""
    return np.einsum("ij,ij->i", kernel_weights(ngbdist, kernel_radius),
                     np.asarray(f, np.float64)[np.asarray(ngb)])


def kernel_variance(ngb, f):
    """np.std(f[ngb], axis=1) (Meshoid.py:352-367), float64."""
    return np.std(np.asarray(f, np.float64)[np.asarray(ngb)], axis=1)


def integrate(vol, f):
    """einsum("i,i...->...", vol, f) over the masked rows (Meshoid.py:333-350)."""
    return np.einsum("i,i...->...", np.asarray(vol, np.float64), np.asarray(f, np.float64))


def num_derivs(dim, order):
    return dim if order == 1 else {1: 2, 2: 5, 3: 9}[dim]


def gradient_weights(pos, ngb, kernel_radius, indices, boxsize=None, weighted=True, order=1,
                     return_status=False):
    pos = _pos2d(pos)
    n, dim = pos.shape
    ngb = np.ascontiguousarray(ngb, dtype=np.int64)
    nq, k = ngb.shape
    h = np.ascontiguousarray(kernel_radius, dtype=np.float64)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    nd = num_derivs(dim, order)
    out = np.empty((nq, nd, k), dtype=np.float64)
    st = np.zeros(nq, dtype=np.int32)
    lib().orc_gradient_weights(pos, n, dim, ngb, nq, k, h, idx,
                               float(boxsize) if boxsize else -1.0, int(bool(weighted)),
                               int(order), out, st)
    return (out, st) if return_status else out


def apply_weights(G, ngb, indices, f, row0=0, nrows=None):
    """out[i, ..., a] = sum_j G[i,row0+a,j] (f[ngb[i,j]] - f[indices[i]])."""
    G = np.ascontiguousarray(G, dtype=np.float64)
    nq, nd, k = G.shape
    nrows = nd - row0 if nrows is None else nrows
    f = np.asarray(f, dtype=np.float64)
    ngb = np.asarray(ngb)
    if ngb.size and (ngb.max() >= f.shape[0] or np.max(indices) >= f.shape[0]):
        # np.take / fancy indexing in the reference raise IndexError here
        raise IndexError("neighbour index out of range for f")
    tail = f.shape[1:]
    f2 = np.ascontiguousarray(f.reshape(f.shape[0], -1))
    K = f2.shape[1]
    out = np.empty((nq, K, nrows), dtype=np.float64)
    lib().orc_apply_weights(G, nq, nd, k, int(row0), int(nrows),
                            np.ascontiguousarray(ngb, dtype=np.int64),
                            np.ascontiguousarray(indices, dtype=np.int64), f2, K, out)
    return out.reshape((nq,) + tail + (nrows,))


def kernel2d(q):
    """The normalised 2D cubic spline, 40/(7 pi) (1 - 6 q^2 (1 - q)) for
    q <= 1/2, 40/(7 pi) 2 (1 - q)^3 for 1/2 < q <= 1 (kernel_density.py:60-68;
    callers only pass q <= 1)."""
    q = np.asarray(q, dtype=np.float64)
    a = 1.0 - q
    return np.where(q <= 0.5, 1.0 - 6.0 * q * q * a, 2.0 * a * a * a) * 1.8189136353359467


def grid_surface_density(f, x, h, center, size, res=100, conservative=False):
    """GridSurfaceDensity(f, x, h, center, size, res, conservative=...)
    (grid_deposition.py:123-295).  The parallel path only splits the
    particles over threads and sums their grids, so the serial cores are
    restated: grid points g*dx, dx = size/(res-1), on x[:, :2] - center +
    size/2; each particle deposits kernel2d(r/h) f/h^2 on the points within
    h (index range int((x - h)/dx + 1) .. int((x + h)/dx), clipped);
    conservative: h >= dx, weights divided by their sum, grid / dx^2."""
    f = np.asarray(f, dtype=np.float64).reshape(-1)
    x2d = np.asarray(x, dtype=np.float64)[:, :2] - np.asarray(center, dtype=np.float64)[:2] + size / 2
    h = np.asarray(h, dtype=np.float64).reshape(-1)
    dx = size / (res - 1)
    grid = np.zeros((res, res))
    for i in range(len(x2d)):
        hs = max(h[i], dx) if conservative else h[i]
        x0, x1 = x2d[i]
        gx0, gx1 = max(int((x0 - hs) / dx + 1), 0), min(int((x0 + hs) / dx), res - 1)
        gy0, gy1 = max(int((x1 - hs) / dx + 1), 0), min(int((x1 + hs) / dx), res - 1)
        if gx1 < gx0 or gy1 < gy0:
            continue
        gx = np.arange(gx0, gx1 + 1)
        gy = np.arange(gy0, gy1 + 1)
        r = np.sqrt(((x0 - gx * dx) ** 2)[:, None] + ((x1 - gy * dx) ** 2)[None, :])
        k = np.where(r <= hs, kernel2d(np.minimum(r / hs, 1.0)), 0.0)
        if conservative:
            tot = k.sum()
            if tot == 0:
                continue
            grid[gx0:gx1 + 1, gy0:gy1 + 1] += k * f[i] / tot
        else:
            grid[gx0:gx1 + 1, gy0:gy1 + 1] += k * (f[i] / hs / hs)
    return grid / (dx * dx) if conservative else grid


def kernel3d(q):
    """The normalised 3D cubic spline, 8/pi (1 - 6 q^2 (1 - q)) for q <= 1/2,
    8/pi 2 (1 - q)^3 for 1/2 < q <= 1 (kernel_density.py:71-79), and 0 beyond
    q = 1, where the reference leaves the value unassigned."""
    q = np.asarray(q, dtype=np.float64)
    a = 1.0 - q
    k = np.where(q <= 0.5, 1.0 - 6.0 * q * q * a, 2.0 * a * a * a) * 2.546479089470325
    return np.where(q <= 1.0, k, 0.0)


def grid_density(f, x, h, center, size, res=100, box_size=-1.0):
    """GridDensity (grid_deposition.py:534-603): grid points at
    center - size/2 + dx/2 + g dx, dx = size/(res-1); each particle spreads f
    over its cube footprint (index range int((x - h)/dx + 1) ..
    int((x + h)/dx), clipped unless box_size >= 0, where indices wrap mod
    res) with weights kernel3d(r/h) / (their sum); the grid is divided by dx^3."""
    f = np.asarray(f, dtype=np.float64).reshape(-1)
    h = np.asarray(h, dtype=np.float64).reshape(-1)
    dx = size / (res - 1)
    x3d = np.asarray(x, dtype=np.float64) - np.asarray(center, dtype=np.float64) + size / 2 - dx / 2
    periodic = box_size is not None and box_size >= 0
    grid = np.zeros((res, res, res))
    for i in range(len(x3d)):
        hs = h[i]
        ranges = []
        for c in range(3):
            lo, hi = int((x3d[i, c] - hs) / dx + 1), int((x3d[i, c] + hs) / dx)
            if not periodic:
                lo, hi = max(lo, 0), min(hi, res - 1)
            ranges.append(np.arange(lo, hi + 1))
        if any(len(r) == 0 for r in ranges):
            continue
        d2 = [(x3d[i, c] - ranges[c] * dx) ** 2 for c in range(3)]
        r = np.sqrt(d2[0][:, None, None] + d2[1][None, :, None] + d2[2][None, None, :])
        k = kernel3d(r / hs)
        tot = k.sum()
        if tot > 0:
            gx, gy, gz = np.meshgrid(ranges[0] % res, ranges[1] % res, ranges[2] % res, indexing="ij")
            np.add.at(grid, (gx, gy, gz), f[i] * k / tot)
    return grid / (dx * dx * dx)


def radtransfer(j, m, kappa, x, h, gridres, L, center=0, i0=0):
    """radtransfer_singlethread (radiation/radtransfer.py:79-177): the
    particles farthest first (largest z first); each changes the pixels of its
    footprint (r <= max(h, dx)) in every band by tau = kappa kernel2d(r/h)
    m / h^2, I <- exp(-tau) I - expm1(-tau) j / (4 pi kappa + 1e-100)."""
    x = np.asarray(x, dtype=np.float64) - np.broadcast_to(np.asarray(center, dtype=np.float64), (3,))
    j = np.asarray(j, dtype=np.float64)
    if j.ndim == 1:
        j = j[:, None]
    nf = j.shape[1]
    kappa = np.asarray(kappa, dtype=np.float64)
    if kappa.ndim == 1:
        kappa = np.repeat(kappa[:, None], nf, axis=1)
    m = np.asarray(m, dtype=np.float64).reshape(-1)
    h = np.asarray(h, dtype=np.float64).reshape(-1)
    order = np.argsort(-x[:, 2], kind="stable")
    dx = L / (gridres - 1)
    intensity = np.zeros((gridres, gridres, nf)) + i0
    for i in order:
        xs = x[i] + L / 2
        hs = max(h[i], dx)
        if hs == 0 or m[i] == 0:
            continue
        src = j[i] / (4 * np.pi * kappa[i] + 1e-100)
        mh2 = m[i] / hs**2
        gx0, gx1 = max(int((xs[0] - hs) / dx + 1), 0), min(int((xs[0] + hs) / dx), gridres - 1)
        gy0, gy1 = max(int((xs[1] - hs) / dx + 1), 0), min(int((xs[1] + hs) / dx), gridres - 1)
        if gx1 < gx0 or gy1 < gy0:
            continue
        gx = np.arange(gx0, gx1 + 1)
        gy = np.arange(gy0, gy1 + 1)
        r = ((xs[0] - gx * dx) ** 2)[:, None] + ((xs[1] - gy * dx) ** 2)[None, :]
        inside = r <= hs * hs
        k = np.where(inside, kernel2d(np.minimum(np.sqrt(r) / hs, 1.0)), 0.0)
        tau = kappa[i][None, None, :] * k[:, :, None] * mh2
        blk = intensity[gx0:gx1 + 1, gy0:gy1 + 1]
        blk[...] = np.where(inside[:, :, None], np.exp(-tau) * blk - np.expm1(-tau) * src[None, None, :], blk)
    return intensity


class OracleMeshoid:
    """FP64 restatement of the reference ``Meshoid`` hot-path state machine
    (``src/meshoid/Meshoid.py:15-170,237-331``), including its quirks:
    ``density`` multiplies ``self.m[self.particle_mask]`` (Meshoid.py:169) and
    ``Curl`` returns (dvy/dz - dvz/dy, dvx/dz - dvz/dx, dvx/dy - dvy/dx)
    (Meshoid.py:311-315)."""

    def __init__(self, pos, m=None, kernel_radius=None, des_ngb=None, boxsize=None,
                 particle_mask=None):
        pos = np.asarray(pos)
        if pos.ndim == 1:
            pos = pos[:, None]
        self.N, self.dim = pos.shape
        if particle_mask is None:
            self.particle_mask = np.arange(self.N)
        elif np.array(particle_mask).dtype == np.dtype("bool"):
            self.particle_mask = np.arange(len(particle_mask))[particle_mask]
        else:
            self.particle_mask = np.asarray(particle_mask)
        self.Nmask = len(self.particle_mask)
        self.des_ngb = des_ngb if des_ngb is not None else {1: 4, 2: 20, 3: 32}[self.dim]
        self.volnorm = {1: 2.0, 2: np.pi, 3: 4 * np.pi / 3}[self.dim]
        self.boxsize = boxsize
        self.pos = pos
        if m is None:
            m = np.repeat(1.0, self.N)
        self.m = np.asarray(m)[self.particle_mask]
        self.ngb = None
        self.kernel_radius = kernel_radius
        self.dweights = self.d2weights = self.dweights_3rdorder = None
        if kernel_radius is None:
            self.TreeUpdate()
        else:
            self.vol = self.volnorm * self.kernel_radius**self.dim / self.des_ngb
            self.density = self.m / self.vol

    def TreeUpdate(self):
        self.ngbdist, self.ngb = knn(self.pos, self.des_ngb, self.boxsize,
                                     queries=np.asarray(self.pos)[self.particle_mask])
        self.kernel_radius = hsml_iter(self.ngbdist, error_norm=1e-13, dim=self.dim)
        self.density = self.des_ngb * self.m[self.particle_mask] / (
            self.volnorm * self.kernel_radius**self.dim)
        self.vol = self.m[self.particle_mask] / self.density

    def ComputeDWeights(self, order=1, weighted=True):
        w = gradient_weights(self.pos, self.ngb, self.kernel_radius, self.particle_mask,
                             boxsize=self.boxsize, weighted=weighted, order=order)
        if order == 1:
            self.dweights = w
        else:
            self.d2weights, self.dweights_3rdorder = w[:, self.dim:, :], w[:, :self.dim, :]

    def D(self, f, order=1, weighted=True):
        if self.ngb is None:
            self.TreeUpdate()
        if order == 1:
            if self.dweights is None:
                self.ComputeDWeights(1, weighted)
            G = self.dweights
        else:
            if self.dweights_3rdorder is None:
                self.ComputeDWeights(2, weighted)
            G = self.dweights_3rdorder
        return apply_weights(G, self.ngb, self.particle_mask, f)

    def D2(self, f, weighted=True):
        if self.ngb is None:
            self.TreeUpdate()
        if self.d2weights is None:
            self.ComputeDWeights(2, weighted)
        return apply_weights(self.d2weights, self.ngb, self.particle_mask, f)

    def Curl(self, v):
        dv = self.D(v[self.particle_mask])
        return np.c_[dv[:, 1, 2] - dv[:, 2, 1], dv[:, 0, 2] - dv[:, 2, 0],
                     dv[:, 0, 1] - dv[:, 1, 0]]

    def Div(self, v):
        dv = self.D(v[self.particle_mask])
        return dv[:, 0, 0] + dv[:, 1, 1] + dv[:, 2, 2]

    # ----------------------------------------------------- reconstruction
    def Reconstruct(self, f, target_points, order=1):
        """Meshoid.py:383-440.  f_target starts as f at the nearest particle
        (in f's dtype) and each term is added in place, so float32 f rounds
        after the first-order term, the pure second-order terms and each mixed
        term."""
        tp = _pos2d(target_points)
        if self.boxsize:
            tp_q = np.mod(tp, self.boxsize)
        else:
            tp_q = tp
        nearest = knn(self.pos, 1, self.boxsize, queries=tp_q)[1][:, 0]
        uniq, inv = np.unique(nearest, return_inverse=True)
        f = np.asarray(f)
        out = f[nearest].copy()
        if order == 0:
            return out
        dx = tp - np.asarray(self.pos, dtype=np.float64)[nearest]
        if self.boxsize:
            L = float(self.boxsize)
            far = np.abs(dx) > L / 2
            dx = np.where(far, -np.sign(dx) * (L - np.abs(dx)), dx)
        if not np.array_equal(self.particle_mask, uniq):
            self.particle_mask = uniq
            self.Nmask = len(uniq)
            self.TreeUpdate()
            self.dweights = self.d2weights = self.dweights_3rdorder = None
        g = self.D(f, order=order)[inv]  # (nt, ..., dim)
        lin = sum(dx[:, d].reshape((-1,) + (1,) * (g.ndim - 2)) * g[..., d] for d in range(self.dim))
        out += lin  # float32 f: one rounding of (f + lin) computed in float64
        if order == 1:
            return out
        h = self.D2(f)[inv]  # (nt, ..., 6): xx, yy, zz, xy, yz, zx
        col = lambda a: a.reshape((-1,) + (1,) * (h.ndim - 2))
        pure = sum(col(dx[:, d] * dx[:, d]) * h[..., d] for d in range(3))
        out += 0.5 * pure
        for d in range(3):
            out += col(dx[:, d] * dx[:, (d + 1) % 3]) * h[..., 3 + d]
        return out

    def Slice(self, f, size, center, res=128, order=1, plane="z"):
        """Meshoid.py:442-502 (size and center given)."""
        ax = np.linspace(-size / 2, size / 2, res)
        x, y = np.meshgrid(ax, ax, indexing="ij")
        z = np.zeros(res * res)
        cols = {"x": (z, x.ravel(), y.ravel()), "y": (x.ravel(), z, y.ravel())}.get(
            plane, (x.ravel(), y.ravel(), z))
        grid = np.stack(cols, axis=1) + center
        return self.Reconstruct(f, grid, order).reshape((res, res) + np.shape(f)[1:])

    def InterpToGrid(self, f, size, center, res=128, order=1):
        """Meshoid.py:504-555 (size and center given): cell centres."""
        e = np.linspace(-size / 2, size / 2, res + 1)
        c = 0.5 * (e[1:] + e[:-1])
        X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
        grid = np.stack((X.ravel(), Y.ravel(), Z.ravel()), axis=1) + center
        return self.Reconstruct(f, grid, order).reshape((res, res, res) + np.shape(f)[1:])

    def SurfaceDensity(self, f=None, size=None, center=None, res=128, smooth_fac=1.0, conservative=False):
        """Meshoid.py:595-643 (size and center given)."""
        if f is None:
            f = self.m
        h = np.clip(smooth_fac * self.kernel_radius, 2 * size / res, 1e100)
        return grid_surface_density(f, self.pos, h, center, size, res, conservative)

    def DepositToGrid(self, f, size, center, res=128):
        """Meshoid.py:557-593 (size and center given)."""
        h = np.clip(self.kernel_radius, size / (res - 1), 1e100)
        return grid_density(f, self.pos, h, center, size, res, -1.0 if self.boxsize is None else self.boxsize)
