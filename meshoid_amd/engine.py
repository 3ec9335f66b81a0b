"""Device-side engine: thin Python objects over the C ABI handles.

PyTorch provides device memory and the stream; every computation happens in
libmeshoid_hip.so.  All tensors passed here must live on the same CUDA (HIP)
device and be contiguous.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import lib, raise_for

_CTX: dict[int, ctypes.c_void_p] = {}


def _dev_index(device) -> int:
    d = torch.device(device) if device is not None else torch.device("cuda")
    if d.type != "cuda":
        raise ValueError(f"meshoid_amd runs on a HIP device, got {d}")
    return d.index if d.index is not None else torch.cuda.current_device()


def context(device=None) -> ctypes.c_void_p:
    """Per-device context bound to torch's current stream of that device."""
    idx = _dev_index(device)
    ctx = _CTX.get(idx)
    if ctx is None:
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device available: meshoid_amd needs an AMD GPU")
        ctx = ctypes.c_void_p()
        rc = lib.mhip_ctx_create(idx, None, ctypes.byref(ctx))
        raise_for(rc, None, "mhip_ctx_create failed")
        _CTX[idx] = ctx
    lib.mhip_ctx_set_stream(ctx, ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream))
    return ctx


def prof_enable(level=1, device=None):
    """1: HIP events around every library launch region; 2: also kNN work
    counters; 0: off."""
    ctx = context(device)
    raise_for(lib.mhip_prof_enable(ctx, int(level)), ctx)


def prof_read(device=None) -> dict:
    """{region: (launches, total_ms)} since the last read (synchronises)."""
    ctx = context(device)
    buf = ctypes.create_string_buffer(1 << 16)
    raise_for(lib.mhip_prof_read(ctx, buf, len(buf)), ctx)
    out = {}
    for line in buf.value.decode().splitlines():
        name, n, ms = line.split()
        out[name] = (int(n), float(ms))
    return out


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _need(t, dtypes, name):
    if t is None:
        return
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype not in dtypes:
        raise ValueError(f"{name} has dtype {t.dtype}, expected one of {dtypes}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class SpatialIndex:
    """Hilbert-sorted particle set + implicit tree (the cKDTree replacement)."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        raise_for(lib.mhip_index_create(ctypes.byref(self.h)), None, "index create")
        self.pos = None
        self.n = 0
        self.dim = 0
        self.boxsize = None

    def build(self, pos: torch.Tensor, boxsize=None):
        _need(pos, (torch.float32, torch.float64), "pos")
        if pos.dim() == 1:
            pos = pos[:, None]
        n, dim = pos.shape
        ctx = context(pos.device)
        rc = lib.mhip_index_build(ctx, self.h, ptr(pos), int(pos.dtype == torch.float64), n, dim,
                                  float(boxsize) if boxsize else -1.0)
        raise_for(rc, ctx)
        self.pos = pos  # the library reads it for float64 boundary resolution
        self.n, self.dim, self.boxsize = n, dim, boxsize
        return self

    def order(self) -> torch.Tensor:
        """Original row of each sorted position (int64, Hilbert order)."""
        perm = torch.empty(self.n, dtype=torch.int32, device=self.pos.device)
        ctx = context(self.pos.device)
        raise_for(lib.mhip_index_order(ctx, self.h, ptr(perm)), ctx)
        return perm.to(torch.int64)

    def bbox(self):
        """Host copy of the bounding box (lo xyz, hi xyz) as float32."""
        bb = (ctypes.c_float * 6)()
        raise_for(lib.mhip_index_info(self.h, None, None, None, bb))
        return [float(v) for v in bb]

    def __del__(self):
        if getattr(self, "h", None):
            lib.mhip_index_destroy(self.h)
            self.h = None


class NeighborList:
    """Exact k-nearest-neighbour rows held on the device (opaque layout)."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        raise_for(lib.mhip_nlist_create(ctypes.byref(self.h)), None, "nlist create")
        self.nq = 0
        self.k = 0
        self.index = None
        self.keep_dist = False

    def export(self, want_ngb=True, want_dist=True, f64=False):
        """(ngb int32 (nq,k), dist (nq,k)) rows ascending, as cKDTree.query.
        f64=True: float64 distances recomputed from the index's positions (the
        distances cKDTree returns); otherwise the list's float32 distances."""
        dev = self.index.pos.device
        ngb = torch.empty((self.nq, self.k), dtype=torch.int32, device=dev) if want_ngb else None
        dt = torch.float64 if f64 else torch.float32
        dist = torch.empty((self.nq, self.k), dtype=dt, device=dev) if want_dist else None
        ctx = context(dev)
        fn = lib.mhip_nlist_export_f64 if f64 else lib.mhip_nlist_export
        raise_for(fn(ctx, self.h, ptr(ngb), ptr(dist)), ctx)
        return ngb, dist

    def h64(self):
        """float64 smoothing lengths of the rows (nq,), refined over the final
        neighbours; the derivative operators fit with these."""
        out = torch.empty(self.nq, dtype=torch.float64, device=self.index.pos.device)
        ctx = context(out.device)
        raise_for(lib.mhip_nlist_hsml(ctx, self.h, ptr(out)), ctx)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib.mhip_nlist_destroy(self.h)
            self.h = None


def knn_hsml(index: SpatialIndex, k: int, qidx=None, m_rows=None, seed_h=None, keep_dist=True,
             nlist: NeighborList | None = None, out=None, seed_r=None, small_record=False):
    """kNN + smoothing length + density.  Returns dict(h, rho, vol, status, nlist).
    seed_h: Newton seed per row; seed_r: search-radius bound per row (rows where
    it is too small are recomputed without it).  Neither changes the result.
    small_record: test hook, every wavefront collects its members by the
    leaf-record overflow path (results unchanged)."""
    dev = index.pos.device
    _need(qidx, (torch.int64,), "qidx")
    _need(m_rows, (torch.float64,), "m_rows")
    _need(seed_h, (torch.float32,), "seed_h")
    _need(seed_r, (torch.float32,), "seed_r")
    nq = index.n if qidx is None else qidx.numel()
    if qidx is not None and nq and (int(qidx.min()) < 0 or int(qidx.max()) >= index.n):
        raise IndexError("query index out of range")
    if m_rows is not None and m_rows.numel() != nq:
        raise ValueError("m_rows must have one entry per query row")
    for sname, sv in (("seed_h", seed_h), ("seed_r", seed_r)):
        if sv is not None and sv.numel() != nq:
            raise ValueError(f"{sname} must have one entry per query row")
    if out is None:
        out = {name: torch.empty(nq, dtype=torch.float32, device=dev) for name in ("h", "rho", "vol")}
        out["status"] = torch.empty(nq, dtype=torch.int32, device=dev)
    nl = nlist if nlist is not None else NeighborList()
    ctx = context(dev)
    rc = lib.mhip_knn_hsml(ctx, index.h, ptr(qidx), nq, int(k), ptr(m_rows), ptr(seed_h),
                           ptr(seed_r), ptr(out.get("h")), ptr(out.get("rho")), ptr(out.get("vol")),
                           ptr(out.get("status")),
                           (_lib.KNN_KEEP_DIST if keep_dist else 0) | (_lib.KNN_SMALL_RECORD if small_record else 0),
                           nl.h)
    raise_for(rc, ctx)
    nl.nq, nl.k, nl.index, nl.keep_dist = nq, int(k), index, keep_dist
    out["nlist"] = nl
    return out


def nlist_subset(full: NeighborList, qidx: torch.Tensor, m_rows=None, nlist: NeighborList | None = None):
    """The rows qidx of an unmasked list: what knn_hsml(index, k, qidx=qidx,
    m_rows=m_rows) returns, without a new search (a particle's neighbours and
    smoothing length do not depend on the mask)."""
    _need(qidx, (torch.int64,), "qidx")
    _need(m_rows, (torch.float64,), "m_rows")
    index = full.index
    dev = index.pos.device
    nq = qidx.numel()
    if nq and (int(qidx.min()) < 0 or int(qidx.max()) >= index.n):
        raise IndexError("query index out of range")
    if m_rows is not None and m_rows.numel() != nq:
        raise ValueError("m_rows must have one entry per query row")
    out = {name: torch.empty(nq, dtype=torch.float32, device=dev) for name in ("h", "rho", "vol")}
    out["status"] = torch.empty(nq, dtype=torch.int32, device=dev)
    nl = nlist if nlist is not None else NeighborList()
    ctx = context(dev)
    raise_for(lib.mhip_nlist_subset(ctx, full.h, ptr(qidx), nq, ptr(m_rows), ptr(out["h"]), ptr(out["rho"]),
                                    ptr(out["vol"]), ptr(out["status"]), nl.h), ctx)
    nl.nq, nl.k, nl.index, nl.keep_dist = nq, full.k, index, full.keep_dist
    out["nlist"] = nl
    return out


def apply_D(nl: NeighborList, f: torch.Tensor, order=1, weighted=True, want_d1=True,
            want_d2=False, out1=None, out2=None, status=None):
    """Least-squares derivatives of f over the list's rows.

    f: (nf_rows, K) contiguous float32/float64 device tensor.  Returns
    (d1 (nq,K,dim) or None, d2 (nq,K,nd2) or None, status int32 (nq,))."""
    _need(f, (torch.float32, torch.float64), "f")
    if f.dim() == 1:
        f = f[:, None]
    nf, K = f.shape
    dim = nl.index.dim
    nd2 = {1: 1, 2: 3, 3: 6}[dim]
    dev = f.device
    if want_d1 and out1 is None:
        out1 = torch.empty((nl.nq, K, dim), dtype=torch.float32, device=dev)
    if want_d2 and out2 is None:
        out2 = torch.empty((nl.nq, K, nd2), dtype=torch.float32, device=dev)
    if status is None:
        status = torch.empty(nl.nq, dtype=torch.int32, device=dev)
    ctx = context(dev)
    rc = lib.mhip_apply_D(ctx, nl.h, ptr(f), int(f.dtype == torch.float64), nf, K, int(order),
                          int(bool(weighted)), ptr(out1 if want_d1 else None),
                          ptr(out2 if want_d2 else None), ptr(status))
    raise_for(rc, ctx)
    return (out1 if want_d1 else None), (out2 if want_d2 else None), status


def gradient_weights_dev(pos, ngb, h, indices, boxsize=None, weighted=True, order=1):
    """G (nq, nd, k) float64 for row-major int32 ngb (device tensors)."""
    _need(pos, (torch.float32, torch.float64), "pos")
    _need(ngb, (torch.int32,), "ngb")
    _need(h, (torch.float64,), "kernel_radius")
    _need(indices, (torch.int64,), "indices")
    if pos.dim() == 1:
        pos = pos[:, None]
    n, dim = pos.shape
    nq, k = ngb.shape
    if dim not in (1, 2, 3):
        raise KeyError(dim)
    nd = dim if order == 1 else {1: 2, 2: 5, 3: 9}[dim]
    if indices.numel() != nq or h.numel() != nq:
        raise ValueError("ngb, kernel_radius and indices must have the same number of rows")
    if nq and (int(ngb.min()) < 0 or int(ngb.max()) >= n or int(indices.min()) < 0
               or int(indices.max()) >= n):
        raise IndexError("neighbour or particle index out of range")
    G = torch.empty((nq, nd, k), dtype=torch.float64, device=pos.device)
    status = torch.empty(nq, dtype=torch.int32, device=pos.device)
    ctx = context(pos.device)
    rc = lib.mhip_gradient_weights(ctx, ptr(pos), int(pos.dtype == torch.float64), n, dim, ptr(ngb),
                                   nq, k, ptr(h), ptr(indices), float(boxsize) if boxsize else -1.0,
                                   int(bool(weighted)), int(order), ptr(G), ptr(status))
    raise_for(rc, ctx)
    return G, status


def hsml_iter_dev(d, dim=3, error_norm=1e-6):
    _need(d, (torch.float32, torch.float64), "neighbor_dists")
    n, k = d.shape
    h = torch.empty(n, dtype=torch.float64, device=d.device)
    status = torch.empty(n, dtype=torch.int32, device=d.device)
    ctx = context(d.device)
    rc = lib.mhip_hsml_iter(ctx, ptr(d), int(d.dtype == torch.float64), n, k, int(dim),
                            float(error_norm), ptr(h), ptr(status))
    raise_for(rc, ctx)
    return h, status


def kernel_dev(q):
    _need(q, (torch.float32, torch.float64), "q")
    out = torch.empty_like(q)
    ctx = context(q.device)
    raise_for(lib.mhip_kernel(ctx, ptr(q), int(q.dtype == torch.float64), q.numel(), ptr(out)), ctx)
    return out


def kernel_weights_dev(dist: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    """W(d/h) normalised per row (Meshoid.get_kernel_weights), float64 (nq, k)."""
    _need(dist, (torch.float32, torch.float64), "dist")
    _need(h, (torch.float32, torch.float64), "h")
    nq, k = dist.shape
    if h.shape[0] != nq:
        raise ValueError("h must have one entry per row of dist")
    w = torch.empty((nq, k), dtype=torch.float64, device=dist.device)
    ctx = context(dist.device)
    raise_for(lib.mhip_kernel_weights(ctx, ptr(dist), int(dist.dtype == torch.float64), ptr(h),
                                      int(h.dtype == torch.float64), nq, k, ptr(w)), ctx)
    return w


def kernel_average(nl: NeighborList, f: torch.Tensor) -> torch.Tensor:
    """Kernel-weighted neighbour average of a scalar field (Meshoid.KernelAverage):
    float64 (nq,)."""
    _need(f, (torch.float32, torch.float64), "f")
    if f.dim() != 1:
        raise ValueError("KernelAverage takes a scalar field of shape (N,)")
    out = torch.empty(nl.nq, dtype=torch.float64, device=f.device)
    ctx = context(f.device)
    raise_for(lib.mhip_kernel_average(ctx, nl.h, ptr(f), int(f.dtype == torch.float64), f.shape[0],
                                      ptr(out)), ctx)
    return out


def kernel_variance(nl: NeighborList, f: torch.Tensor) -> torch.Tensor:
    """Population standard deviation of f over each row's neighbours
    (Meshoid.KernelVariance): float64 (nq, K) for f of shape (N, K)."""
    _need(f, (torch.float32, torch.float64), "f")
    f2 = f.reshape(f.shape[0], -1).contiguous() if f.dim() != 2 else f
    K = f2.shape[1]
    out = torch.empty((nl.nq, K), dtype=torch.float64, device=f.device)
    ctx = context(f.device)
    raise_for(lib.mhip_kernel_variance(ctx, nl.h, ptr(f2), int(f.dtype == torch.float64), f2.shape[0],
                                       K, ptr(out)), ctx)
    return out


def integrate(vol: torch.Tensor, f: torch.Tensor) -> torch.Tensor:
    """sum_i vol_i f_i per component (Meshoid.Integrate): float64 (K,)."""
    _need(vol, (torch.float32, torch.float64), "vol")
    _need(f, (torch.float32, torch.float64), "f")
    f2 = f.reshape(f.shape[0], -1).contiguous()
    if f2.shape[0] != vol.shape[0]:
        raise ValueError("f and vol must have the same number of rows")
    K = f2.shape[1]
    out = torch.empty(K, dtype=torch.float64, device=f.device)
    ctx = context(f.device)
    raise_for(lib.mhip_integrate(ctx, ptr(vol), int(vol.dtype == torch.float64), ptr(f2),
                                 int(f.dtype == torch.float64), f2.shape[0], K, ptr(out)), ctx)
    return out


def integrate_noalloc(vol: torch.Tensor, f: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
    """integrate() through mhip_integrate_noalloc, which allocates no device
    memory (the caller passes out, float64 (K,)); K <= 64 components."""
    _need(vol, (torch.float32, torch.float64), "vol")
    _need(f, (torch.float32, torch.float64), "f")
    _need(out, (torch.float64,), "out")
    f2 = f.reshape(f.shape[0], -1)
    if f2.shape[0] != vol.shape[0]:
        raise ValueError("f and vol must have the same number of rows")
    K = f2.shape[1]
    if out.numel() != K:
        raise ValueError("out must hold one value per component")
    ctx = context(f.device)
    raise_for(lib.mhip_integrate_noalloc(ctx, ptr(vol), int(vol.dtype == torch.float64), ptr(f2),
                                         int(f.dtype == torch.float64), f2.shape[0], K, ptr(out)), ctx)
    return out


def nearest(index: SpatialIndex, points: torch.Tensor) -> torch.Tensor:
    """Nearest particle of each point, cKDTree.query(points)[1]
    (src/meshoid/Meshoid.py:409): points (nt, dim) float64 device tensor;
    int64 (nt,) original particle indices, exact in float64."""
    _need(points, (torch.float64,), "target_points")
    if points.dim() == 1:
        points = points[:, None]
    nt, dim = points.shape
    if dim != index.dim:
        raise ValueError(f"target_points have dimension {dim}, the particles {index.dim}")
    out = torch.empty(nt, dtype=torch.int64, device=points.device)
    ctx = context(points.device)
    raise_for(lib.mhip_nearest(ctx, index.h, ptr(points), nt, dim, ptr(out)), ctx)
    return out


def unique_inverse(idx: torch.Tensor, n: int):
    """np.unique(idx, return_inverse=True) for int64 indices in [0, n)
    (src/meshoid/Meshoid.py:410): (sorted distinct indices, inverse)."""
    _need(idx, (torch.int64,), "indices")
    nt = idx.numel()
    uniq = torch.empty(max(1, min(n, nt)), dtype=torch.int64, device=idx.device)
    inv = torch.empty(nt, dtype=torch.int64, device=idx.device)
    cnt = ctypes.c_int64()
    ctx = context(idx.device)
    raise_for(lib.mhip_unique_inverse(ctx, ptr(idx), nt, int(n), ptr(uniq) if nt else None,
                                      ptr(inv) if nt else None, ctypes.byref(cnt)), ctx)
    return uniq[: cnt.value], inv


def reconstruct(points, pos, boxsize, nbr, inv, f, d1=None, d2=None, order=1):
    """f_target of Meshoid.Reconstruct (src/meshoid/Meshoid.py:413-440):
    points (nt, dim) float64, pos (n, dim), nbr/inv int64 (nt,), f (n, K)
    float32/float64, d1 (nu, K, dim) / d2 (nu, K, 6) float32 on the unique
    rows; returns (nt, K) in f's dtype."""
    _need(points, (torch.float64,), "target_points")
    _need(pos, (torch.float32, torch.float64), "pos")
    _need(f, (torch.float32, torch.float64), "f")
    _need(nbr, (torch.int64,), "nbr")
    _need(inv, (torch.int64,), "inv")
    _need(d1, (torch.float32,), "d1")
    _need(d2, (torch.float32,), "d2")
    if points.dim() == 1:
        points = points[:, None]
    nt, dim = points.shape
    K = f.shape[1] if f.dim() > 1 else 1
    if order >= 1 and d1 is None:
        raise ValueError("order >= 1 needs first derivatives")
    if order == 2 and d2 is None:
        raise ValueError("order 2 needs second derivatives")
    out = torch.empty((nt, K), dtype=f.dtype, device=f.device)
    ctx = context(f.device)
    raise_for(lib.mhip_reconstruct(ctx, ptr(points), nt, dim, float(boxsize) if boxsize else -1.0, ptr(pos),
                                   int(pos.dtype == torch.float64), ptr(nbr), ptr(inv), ptr(f),
                                   int(f.dtype == torch.float64), K, ptr(d1), ptr(d2), int(order), ptr(out)),
              ctx)
    return out


def surface_density(f, pos, h, center, size, res=128, smooth_fac=1.0, conservative=False, hmin=None):
    """GridSurfaceDensity (src/meshoid/grid_deposition.py:123-295) of f at
    positions pos (n, 2|3) with smoothing lengths clip(smooth_fac h, hmin,
    1e100), hmin defaulting to SurfaceDensity's 2 size/res:
    (res, res) float64 device grid."""
    _need(f, (torch.float32, torch.float64), "f")
    _need(pos, (torch.float32, torch.float64), "pos")
    _need(h, (torch.float32, torch.float64), "h")
    n, dim = pos.shape
    if f.numel() != n or h.numel() != n:
        raise IndexError("f, positions and smoothing lengths must have one entry per particle")
    grid = torch.empty((int(res), int(res)), dtype=torch.float64, device=pos.device)
    ctx = context(pos.device)
    raise_for(lib.mhip_surface_density(ctx, ptr(f), int(f.dtype == torch.float64), ptr(pos),
                                       int(pos.dtype == torch.float64), n, dim, ptr(h), int(h.dtype == torch.float64),
                                       float(smooth_fac), float(2.0 * size / res if hmin is None else hmin),
                                       float(center[0]), float(center[1]), float(size), int(res),
                                       int(bool(conservative)), ptr(grid)), ctx)
    return grid


def grid_density(f, pos, h, center, size, res=128, periodic=False, hmin=0.0):
    """GridDensity (src/meshoid/grid_deposition.py:534-603): (res, res, res)
    float64 device grid of the density of f, smoothing lengths clipped below
    at hmin; periodic wraps the grid indices."""
    _need(f, (torch.float32, torch.float64), "f")
    _need(pos, (torch.float32, torch.float64), "pos")
    _need(h, (torch.float32, torch.float64), "h")
    n = pos.shape[0]
    if pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError("GridDensity needs (N, 3) positions")
    if f.numel() != n or h.numel() != n:
        raise IndexError("f, positions and smoothing lengths must have one entry per particle")
    grid = torch.empty((int(res),) * 3, dtype=torch.float64, device=pos.device)
    c = [float(v) for v in center]
    ctx = context(pos.device)
    raise_for(lib.mhip_grid_density(ctx, ptr(f), int(f.dtype == torch.float64), ptr(pos),
                                    int(pos.dtype == torch.float64), n, ptr(h), int(h.dtype == torch.float64),
                                    float(hmin), c[0], c[1], c[2], float(size), int(res), int(bool(periodic)),
                                    ptr(grid)), ctx)
    return grid
