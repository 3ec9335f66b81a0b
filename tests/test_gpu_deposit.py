"""SurfaceDensity / GridSurfaceDensity (src/meshoid/Meshoid.py:595-643,
src/meshoid/grid_deposition.py:123-295) on the GPU against the oracle
(tests/test_deposit_oracle.py pins it): both modes, the LDS-resident grid
(res <= 128) and the HBM-atomic grid (res > 128), particles off the map,
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
