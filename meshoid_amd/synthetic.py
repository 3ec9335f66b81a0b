"""Seeded synthetic particle sets of the benchmark workloads (NumPy, host).

* :func:`uniform`    -- positions uniform in the unit cube.
* :func:`clustered`  -- mixture of isotropic Gaussian clumps: centres uniform in
  a ``box``-unit cube, sigma log-uniform in [0.05, 5], clump weights from the
  truncated power law dN/dM ~ M^-2 on [1, 1e4], particle counts multinomial;
  optionally a uniform background fraction.  About six decades of density.
* :func:`sin3`       -- f = sin(kx) sin(ky) sin(kz) and its exact gradient.
* :func:`rotation_field` -- v = Omega x (r - c) + N(0, noise^2).
* :func:`grf`        -- Gaussian random field with P(k) ~ k^slope on a periodic
  grid, seeded, trilinearly interpolated to the particles.
"""
from __future__ import annotations

import numpy as np


def uniform(n, dim=3, seed=0, dtype=np.float64):
    return np.random.default_rng(seed).random((n, dim)).astype(dtype)


def clump_params(nclump, seed, box=100.0, sig_lo=0.05, sig_hi=5.0, m_lo=1.0, m_hi=1e4):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, box, (nclump, 3))
    sigma = np.exp(rng.uniform(np.log(sig_lo), np.log(sig_hi), nclump))
    u = rng.random(nclump)
    # inverse CDF of p(M) ~ M^-2 on [m_lo, m_hi]
    mass = 1.0 / (1.0 / m_lo - u * (1.0 / m_lo - 1.0 / m_hi))
    return rng, centers, sigma, mass


def clustered(n, seed=0, nclump=256, box=100.0, background=0.0, dtype=np.float32):
    """(n, 3) positions; the first n_bg = round(background*n) rows are the
    uniform background when background > 0."""
    rng, centers, sigma, mass = clump_params(nclump, seed, box)
    n_bg = int(round(background * n))
    counts = rng.multinomial(n - n_bg, mass / mass.sum())
    out = np.empty((n, 3), dtype=np.float64)
    out[:n_bg] = rng.uniform(0.0, box, (n_bg, 3))
    o = n_bg
    for c, s, m in zip(centers, sigma, counts):
        out[o:o + m] = c + s * rng.standard_normal((m, 3))
        o += m
    return out.astype(dtype)


def sin3(pos, k=2 * np.pi / 10):
    x = np.asarray(pos, dtype=np.float64)
    s = np.sin(k * x)
    c = np.cos(k * x)
    f = s[:, 0] * s[:, 1] * s[:, 2]
    g = np.stack([k * c[:, 0] * s[:, 1] * s[:, 2], k * s[:, 0] * c[:, 1] * s[:, 2],
                  k * s[:, 0] * s[:, 1] * c[:, 2]], axis=1)
    return f, g


def rotation_field(pos, omega=(0.0, 0.0, 1.0), center=50.0, noise=0.1, seed=0):
    x = np.asarray(pos, dtype=np.float64) - center
    v = np.cross(np.asarray(omega, dtype=np.float64), x)
    if noise:
        v = v + noise * np.random.default_rng(seed).standard_normal(v.shape)
    return v


def grf_grid(ngrid=256, box=100.0, slope=-3.0, seed=0):
    """Periodic Gaussian random field on an ngrid^3 grid over [0, box)^3 with
    power spectrum P(k) ~ k^slope (the k = 0 mode removed), unit variance."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((ngrid, ngrid, ngrid))
    wk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(ngrid, d=box / ngrid) * 2 * np.pi
    kz = np.fft.rfftfreq(ngrid, d=box / ngrid) * 2 * np.pi
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    amp = np.zeros_like(k2)
    nz = k2 > 0
    amp[nz] = k2[nz] ** (slope / 4.0)  # sqrt(P(k)) with P ~ |k|^slope
    field = np.fft.irfftn(wk * amp, s=white.shape, axes=(0, 1, 2))
    field -= field.mean()
    return field / field.std()


def trilinear_periodic(grid, pos, box=100.0):
    """Trilinear interpolation of a periodic grid over [0, box)^3 at pos."""
    ng = grid.shape[0]
    u = np.mod(np.asarray(pos, dtype=np.float64), box) * (ng / box)
    i0 = np.floor(u).astype(np.int64)
    t = u - i0
    i0 %= ng
    i1 = (i0 + 1) % ng
    out = np.zeros(len(u))
    for dx in (0, 1):
        wx = t[:, 0] if dx else 1 - t[:, 0]
        ix = i1[:, 0] if dx else i0[:, 0]
        for dy in (0, 1):
            wy = t[:, 1] if dy else 1 - t[:, 1]
            iy = i1[:, 1] if dy else i0[:, 1]
            for dz in (0, 1):
                wz = t[:, 2] if dz else 1 - t[:, 2]
                iz = i1[:, 2] if dz else i0[:, 2]
                out += wx * wy * wz * grid[ix, iy, iz]
    return out


def grf(pos, ngrid=256, box=100.0, slope=-3.0, seed=0):
    """GRF values at the particles (float64)."""
    return trilinear_periodic(grf_grid(ngrid, box, slope, seed), pos, box)


def snapshot_batch(nsnap=64, K=32, seed=0, n_lo=16384, n_hi=1048576, box=100.0):
    """Uneven particle sets for the batched form: N_s = round(exp(U[ln n_lo,
    ln n_hi])), 8-128 clumps each (10% background), des_ngb alternating 32 /
    64, and K float32 fields N(0,1) + a.x + x^T B x per set.  Returns a list
    of (pos float32 (N_s, 3), fields float32 (N_s, K), des_ngb)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nsnap):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        pos = clustered(n, seed=seed * 1000 + s, nclump=int(rng.integers(8, 129)), box=box, background=0.1)
        x = pos.astype(np.float64) / box
        a = rng.standard_normal((3, K))
        B = rng.standard_normal((K, 3, 3))
        # x^T B x = sum_ij B_ij x_i x_j through the 6 quadratic monomials
        q = np.stack([x[:, 0] ** 2, x[:, 1] ** 2, x[:, 2] ** 2, x[:, 0] * x[:, 1], x[:, 1] * x[:, 2],
                      x[:, 0] * x[:, 2]], axis=1)
        C = np.stack([B[:, 0, 0], B[:, 1, 1], B[:, 2, 2], B[:, 0, 1] + B[:, 1, 0], B[:, 1, 2] + B[:, 2, 1],
                      B[:, 0, 2] + B[:, 2, 0]], axis=0)
        f = rng.standard_normal((n, K), dtype=np.float32) + (x @ a + q @ C).astype(np.float32)
        out.append((pos, f, 32 if s % 2 == 0 else 64))
    return out
