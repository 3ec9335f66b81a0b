"""SurfaceDensity / GridSurfaceDensity (src/meshoid/Meshoid.py:595-643,
src/meshoid/grid_deposition.py:123-295) on the GPU against the oracle
(tests/test_deposit_oracle.py pins it): both modes, the LDS-resident grid
(res <= 128) and the binned LDS-window path (res > 128), particles off the map,
footprints large enough for the block-per-particle kernel, float32 and
float64 inputs, 2D and 3D positions.  Float64 sums in another order: maps
agree to 1e-11 of their maximum."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _close(g, go, what):
    g = np.asarray(g, np.float64)
    err = np.abs(g - go).max() / max(np.abs(go).max(), 1e-300)
    say(what, "max rel to map max", err)
    assert err <= 1e-11, what


def test_grid_surface_density_paths(oracle):
    torch = _torch()
    from meshoid_amd import engine

    rng = np.random.default_rng(0)
    n = 6000
    pos = rng.random((n, 3)) * 12 - 1  # part of the particles off the [0, 10] map
    f = rng.random(n) + 0.1
    h = np.exp(rng.normal(-1.5, 0.8, n))
    h[:20] = 4.0  # footprints over 1024 cells at res 200 / 128: the block-per-particle kernel
    c = np.array([5.0, 5.0])
    for res in (32, 128, 200):
        for cons in (False, True):
            for dt in (np.float64, np.float32):
                say("surface density res", res, "conservative", cons, dt.__name__)
                g = engine.surface_density(torch.from_numpy(f.astype(dt)).cuda(),
                                           torch.from_numpy(pos.astype(dt)).cuda(),
                                           torch.from_numpy(h.astype(dt)).cuda(), c, 10.0, res, 1.3, cons)
                hh = np.clip(1.3 * h.astype(dt).astype(np.float64), 2 * 10.0 / res, 1e100)
                go = oracle.grid_surface_density(f.astype(dt), pos.astype(dt), hh, c, 10.0, res, cons)
                _close(g.cpu().numpy(), go, f"res {res} cons {cons} {dt.__name__}")


def test_meshoid_surface_density_2d_and_3d(oracle):
    _torch()
    import meshoid_amd
    from meshoid_amd import synthetic

    pos = synthetic.clustered(40000, seed=3, nclump=8, box=10.0, background=0.2).astype(np.float64)
    for dim in (3, 2):
        p = np.ascontiguousarray(pos[:, :dim])
        M = meshoid_amd.Meshoid(p, des_ngb=20 if dim == 2 else 32)
        O = oracle.OracleMeshoid(p, des_ngb=20 if dim == 2 else 32)
        c = np.array([5.0, 5.0, 5.0][:dim])
        for cons in (False, True):
            say("Meshoid.SurfaceDensity dim", dim, "conservative", cons)
            s = M.SurfaceDensity(size=8.0, center=c, res=96, conservative=cons)
            so = O.SurfaceDensity(size=8.0, center=c, res=96, conservative=cons)
            assert s.shape == (96, 96) and s.dtype == np.float64
            # h on the GPU is float32-exact to 1e-6 of the oracle's: compare to 1e-5
            err = np.abs(s - so).max() / np.abs(so).max()
            say("  max rel to map max", err)
            assert err <= 1e-5
    # default size / center (the meshoid's frame) and mass weighting
    M = meshoid_amd.Meshoid(pos, np.full(len(pos), 2.0), des_ngb=32)
    s = M.SurfaceDensity(res=64, conservative=True)
    dx = M.L / 63
    inside = np.all(np.abs(pos[:, :2] - M.center[:2]) < M.L / 2 - 2 * M.L / 64, axis=1)
    # the conservative map holds at least the mass of the particles well inside
    assert s.sum() * dx * dx >= 2.0 * inside.sum() * (1 - 1e-9)


def test_grid_density_matches_oracle(oracle):
    """GridDensity / DepositToGrid (grid_deposition.py:534-603,
    Meshoid.py:557-593): open and periodic grids, float32 and float64, large
    footprints on the block-per-particle kernel."""
    torch = _torch()
    import meshoid_amd
    from meshoid_amd import engine

    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.random((n, 3)) * 11 - 0.5
    f = rng.random(n) + 0.1
    h = np.exp(rng.normal(-1.0, 0.6, n))
    h[:10] = 3.5  # footprints over 512 points at res 24 / 40
    c = np.array([5.0, 5.0, 5.0])
    for res in (24, 40):
        for periodic in (False, True):
            for dt in (np.float64, np.float32):
                say("grid density res", res, "periodic", periodic, dt.__name__)
                g = engine.grid_density(torch.from_numpy(f.astype(dt)).cuda(), torch.from_numpy(pos.astype(dt)).cuda(),
                                        torch.from_numpy(h.astype(dt)).cuda(), c, 10.0, res, periodic, 0.0)
                go = oracle.grid_density(f.astype(dt), pos.astype(dt), h.astype(dt), c, 10.0, res,
                                         10.0 if periodic else -1.0)
                _close(g.cpu().numpy(), go, f"grid density res {res} periodic {periodic} {dt.__name__}")
    # the module function and the Meshoid method (periodic box, as the reference needs)
    g = meshoid_amd.GridDensity(f, pos, h, c, 10.0, res=24, box_size=10.0)
    _close(g, oracle.grid_density(f, pos, h, c, 10.0, 24, 10.0), "GridDensity module function")
    pb = np.mod(rng.random((20000, 3)) * 10.0, 10.0)
    M = meshoid_amd.Meshoid(pb, des_ngb=32, boxsize=10.0)
    O = oracle.OracleMeshoid(pb, des_ngb=32, boxsize=10.0)
    m = rng.random(20000)
    d = M.DepositToGrid(m, size=10.0, center=c, res=32)
    do = O.DepositToGrid(m, 10.0, c, res=32)
    err = np.abs(d - do).max() / np.abs(do).max()
    say("DepositToGrid max rel to map max", err)
    assert d.shape == (32, 32, 32) and err <= 1e-5  # the float32 h
    dx = 10.0 / 31
    assert abs(d.sum() * dx**3 / m.sum() - 1.0) < 1e-9  # periodic: every particle's mass on the grid


def test_grid_surface_density_module_function(oracle):
    _torch()
    import meshoid_amd

    rng = np.random.default_rng(6)
    pos = rng.random((4000, 3)) * 10
    f = rng.random(4000)
    h = rng.random(4000) * 0.3 + 0.01  # no clip: small h deposit on few or no points
    for cons in (False, True):
        g = meshoid_amd.GridSurfaceDensity(f, pos, h, [5.0, 5.0], 10.0, res=64, conservative=cons)
        _close(g, oracle.grid_surface_density(f, pos, h, [5.0, 5.0], 10.0, 64, cons), f"module cons {cons}")


def test_binned_deposition_clustered(oracle):
    """The binned path on clustered particles: bricks whose runs span several
    blocks and go through the LDS window, sparse background bricks on HBM
    atomics, footprints reaching the window's far edge, periodic wrap."""
    torch = _torch()
    from meshoid_amd import engine, synthetic

    pos = synthetic.clustered(100000, seed=7, nclump=16, box=10.0, background=0.1).astype(np.float64)
    rng = np.random.default_rng(1)
    h = np.exp(rng.normal(-3.0, 0.8, len(pos)))
    h[rng.random(len(pos)) < 0.01] = 0.4  # spans of ~33 map cells at res 400, ~6 volume cells at res 64
    f = rng.random(len(pos)) + 0.1
    td = lambda a: torch.from_numpy(a).cuda()
    c2 = np.array([5.0, 5.0])
    for cons in (False, True):
        say("binned surface density res 400 conservative", cons)
        g = engine.surface_density(td(f), td(pos), td(h), c2, 10.0, 400, 1.0, cons)
        go = oracle.grid_surface_density(f, pos, np.clip(h, 2 * 10.0 / 400, 1e100), c2, 10.0, 400, cons)
        _close(g.cpu().numpy(), go, f"binned res 400 cons {cons}")
    c3 = np.array([5.0, 5.0, 5.0])
    dx = 10.0 / 63
    for periodic in (False, True):
        say("binned grid density res 64 periodic", periodic)
        g = engine.grid_density(td(f), td(pos), td(h), c3, 10.0, 64, periodic, dx)
        go = oracle.grid_density(f, pos, np.maximum(h, dx), c3, 10.0, 64, 10.0 if periodic else -1.0)
        _close(g.cpu().numpy(), go, f"binned grid density periodic {periodic}")


@pytest.mark.parametrize("cons", [False, True])
def test_surface_density_map_beyond_int32_cells(oracle, cons):
    """A 46500^2 map (more cells than int32 indexes) with footprints on the
    block-per-particle path near the far corner: every deposit lands on its
    own cells.  Each footprint is compared with the oracle's map of a 200^2
    window at the same grid spacing (dx = 1)."""
    torch = _torch()
    from meshoid_amd import engine

    R = 46500
    S = float(R - 1)  # grid spacing 1
    pts = np.array([[46380.3, 46370.7], [46390.6, 120.2], [150.4, 46385.9]])
    h = np.array([40.0, 33.0, 37.5])  # footprints of ~5000 cells
    f = np.array([1.0, 2.0, 0.5])
    pos = np.concatenate([pts, np.zeros((3, 1))], axis=1)
    say("surface density res", R, "conservative", cons)
    g = engine.surface_density(torch.from_numpy(f).cuda(), torch.from_numpy(pos).cuda(),
                               torch.from_numpy(h).cuda(), (S / 2, S / 2), S, R, 1.0, cons, hmin=0.0)
    total = float(g.sum())
    win = 0.0
    for i in range(3):
        g0 = np.floor(pts[i] - 100).astype(int)
        sub = g[g0[0]:g0[0] + 200, g0[1]:g0[1] + 200].cpu().numpy()
        win += sub.sum()
        # the same footprint on a 200^2 map whose first point is cell g0
        c_s = g0 + 99.5
        go = oracle.grid_surface_density(f[i:i + 1], pos[i:i + 1], h[i:i + 1], c_s, 199.0, 200, cons)
        _close(sub, go, f"footprint {i}")
    del g
    torch.cuda.empty_cache()
    np.testing.assert_allclose(win, total, rtol=1e-12)
