"""CPU checks of the oracle's GridSurfaceDensity restatement
(src/meshoid/grid_deposition.py:123-295, kernel2d kernel_density.py:60-68)
against what the mathematics fixes: the 2D kernel integrates to one, the
conservative mode deposits exactly f, a particle on a grid point puts
kernel2d(0) f / h^2 there, and a centred particle gives a symmetric map."""
import numpy as np

from oracle import oracle as orc


def test_kernel2d_normalised_and_piecewise():
    # integral of kernel2d(r) over the unit disc = 1 (radial quadrature)
    r = (np.arange(200000) + 0.5) / 200000
    assert abs(np.sum(orc.kernel2d(r) * 2 * np.pi * r) / 200000 - 1.0) < 1e-9
    c = 40 / (7 * np.pi)
    assert np.isclose(orc.kernel2d(0.0), c) and np.isclose(orc.kernel2d(1.0), 0.0)
    assert np.isclose(orc.kernel2d(0.5), c * 0.25)  # both branches meet: 1 - 6/8 = 2/8
    assert np.isclose(orc.kernel2d(0.75), c * 2 * 0.25**3)


def test_non_conservative_integral_and_point_value():
    # one particle, h many grid spacings: sum(grid) dx^2 -> f
    size, res = 10.0, 201
    dx = size / (res - 1)
    g = orc.grid_surface_density([2.5], [[0.3, -0.2, 7.0]], [1.7], [0.0, 0.0], size, res)
    assert abs(g.sum() * dx * dx / 2.5 - 1.0) < 2e-4
    # particle on grid point (100, 100): the value there is kernel2d(0) f / h^2
    g = orc.grid_surface_density([3.0], [[0.0, 0.0]], [0.8], [0.0, 0.0], size, res)
    assert np.isclose(g[100, 100], 40 / (7 * np.pi) * 3.0 / 0.64, rtol=1e-14)
    # ... and is symmetric under the square's symmetries
    # (the flip up to the rounding of 5 - (100 +/- k) dx)
    assert np.array_equal(g, g.T) and np.allclose(g, g[::-1, ::-1], rtol=0, atol=1e-13 * g.max())
    # a point exactly h away receives nothing (W(1) = 0): g[100, 108] with dx = 0.05 is 0.4 away
    g = orc.grid_surface_density([1.0], [[0.0, 0.0]], [0.4], [0.0, 0.0], size, res)
    assert g[100, 108] == 0.0 and g[100, 107] > 0.0


def test_conservative_deposits_f_exactly():
    rng = np.random.default_rng(0)
    n = 300
    x = rng.random((n, 3)) * 6 + 2  # inside [2, 8]: every footprint on the grid
    f = rng.random(n) + 0.1
    h = rng.random(n) * 0.8 + 0.05  # some below dx: clamped to dx
    size, res = 10.0, 64
    dx = size / (res - 1)
    g = orc.grid_surface_density(f, x, h, [5.0, 5.0], size, res, conservative=True)
    assert abs(g.sum() * dx * dx - f.sum()) < 1e-12 * f.sum()
    # the non-conservative map only approximates the total (discrete sum of the kernel)
    g2 = orc.grid_surface_density(f, x, np.maximum(h, 2 * size / res), [5.0, 5.0], size, res)
    assert abs(g2.sum() * dx * dx / f.sum() - 1.0) < 0.1


def test_surface_density_clip():
    rng = np.random.default_rng(1)
    pos = rng.random((500, 3))
    O = orc.OracleMeshoid(pos, des_ngb=16)
    res, size = 32, 1.0
    s = O.SurfaceDensity(size=size, center=[0.5, 0.5, 0.5], res=res, smooth_fac=0.01)
    # every h clipped up to 2 size / res
    want = orc.grid_surface_density(O.m, pos, np.full(500, 2 * size / res), [0.5, 0.5], size, res)
    assert np.array_equal(s, want)


def test_kernel3d_normalised():
    r = (np.arange(200000) + 0.5) / 200000
    assert abs(np.sum(orc.kernel3d(r) * 4 * np.pi * r * r) / 200000 - 1.0) < 1e-9
    assert orc.kernel3d(1.5) == 0.0 and np.isclose(orc.kernel3d(0.0), 8 / np.pi)


def test_grid_density_conserves_and_points():
    rng = np.random.default_rng(2)
    n = 200
    size, res = 10.0, 24
    dx = size / (res - 1)
    x = rng.random((n, 3)) * 6 + 2
    f = rng.random(n) + 0.1
    h = rng.random(n) * 1.5 + dx
    g = orc.grid_density(f, x, h, [5.0, 5.0, 5.0], size, res)
    assert abs(g.sum() * dx**3 - f.sum()) < 1e-12 * f.sum()
    # on grid point (10, 11, 12) with h = dx: the six neighbours sit at q = 1
    # (weight 0), so all of f lands on that point
    p = np.array([[10.0, 11.0, 12.0]]) * dx + 5.0 - size / 2 + dx / 2
    g = orc.grid_density([2.0], p, [dx], [5.0, 5.0, 5.0], size, res)
    assert np.isclose(g[10, 11, 12], 2.0 / dx**3, rtol=1e-12) and np.count_nonzero(g > 1e-12 * g.max()) == 1
    # periodic: a particle at the corner wraps onto the far faces and keeps its mass
    g = orc.grid_density([1.0], [[0.0, 0.0, 0.0]], [2.5 * dx], [5.0, 5.0, 5.0], size, res, box_size=size)
    assert abs(g.sum() * dx**3 - 1.0) < 1e-12 and g[-1, -1, -1] > 0 and g[0, 0, 0] > 0
    # clipped (open) grid: the weights are normalised over the points on the
    # grid, so the mass stays whole, all of it in the corner
    g2 = orc.grid_density([1.0], [[0.0, 0.0, 0.0]], [2.5 * dx], [5.0, 5.0, 5.0], size, res)
    assert abs(g2.sum() * dx**3 - 1.0) < 1e-12 and g2[-1, -1, -1] == 0.0
    assert np.count_nonzero(g2[:3, :3, :3]) == np.count_nonzero(g2)
