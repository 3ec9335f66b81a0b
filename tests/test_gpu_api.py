"""The rest of the reference's Meshoid surface on the GPU against the FP64
oracle: the kernel_radius= / hsml= constructor branch
(src/meshoid/Meshoid.py:94-105), the ComputeDWeights split into dweights,
d2weights and dweights_3rdorder (:110-140), the accessors NeighborDistance,
Volume and NearestNeighbors (:179-211), and the kernel weighting of
gradient_weights (src/meshoid/derivatives.py:50-51) on a stencil with a
closed-form answer."""
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


def rel_rows(a, b):
    a = np.asarray(a, np.float64).reshape(len(a), -1)
    b = np.asarray(b, np.float64).reshape(len(b), -1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


@pytest.mark.parametrize("kw", ["kernel_radius", "hsml", "positional"])
def test_given_kernel_radius_branch(oracle, kw):
    _torch()
    import meshoid_amd

    say("kernel_radius branch", kw)
    rng = np.random.default_rng(4)
    pos = rng.random((8000, 3))
    m = rng.random(8000) + 0.5
    h = 0.05 + 0.02 * rng.random(8000)
    if kw == "positional":
        M = meshoid_amd.Meshoid(pos, m, h)
    else:
        M = meshoid_amd.Meshoid(pos, m, **{kw: h})
    O = oracle.OracleMeshoid(pos, m, kernel_radius=h)
    # Meshoid.py:104-105: vol = V_d h^d / des_ngb, density = m / vol (float64)
    vol = 4 * np.pi / 3 * h**3 / 32
    np.testing.assert_allclose(M.Volume(), vol, rtol=1e-15)
    np.testing.assert_allclose(M.Density(), m / vol, rtol=1e-15)
    np.testing.assert_array_equal(M.SmoothingLength(), h)
    np.testing.assert_allclose(M.vol, O.vol, rtol=1e-15)
    # ngbdist is never set on this branch (Meshoid.py:94-105)
    with pytest.raises(AttributeError):
        M.NeighborDistance()
    # D runs TreeUpdate, which replaces the given h by the adaptive one
    f = pos[:, 0] ** 2 - pos[:, 1] * pos[:, 2]
    g = M.D(f)
    g_o = O.D(f)
    assert rel_rows(g, g_o).max() <= 1e-5
    np.testing.assert_allclose(M.kernel_radius, O.kernel_radius, rtol=1e-6)
    assert np.abs(M.kernel_radius - h).max() > 1e-3  # no longer the given radii


def test_kernel_radius_branch_nearest_neighbors_runs_treeupdate(oracle):
    _torch()
    import meshoid_amd

    rng = np.random.default_rng(6)
    pos = rng.random((5000, 3))
    M = meshoid_amd.Meshoid(pos, kernel_radius=np.full(5000, 0.1))
    ngb = M.NearestNeighbors()  # Meshoid.py:197-198: TreeUpdate when ngb is None
    d_o, n_o = oracle.knn(pos, 32)
    assert ngb.shape == (5000, 32) and ngb.dtype == np.int64
    assert all(set(a) == set(b) for a, b in zip(ngb.tolist(), n_o.tolist()))
    np.testing.assert_allclose(M.NeighborDistance(), d_o, rtol=1e-15, atol=0)


@pytest.mark.parametrize("dtype,box", [(np.float64, None), (np.float32, None), (np.float32, 1.0)])
def test_accessors_match_oracle(oracle, dtype, box):
    _torch()
    import meshoid_amd

    say("accessors", dtype.__name__, box)
    rng = np.random.default_rng(9)
    pos = rng.random((12000, 3)).astype(dtype)
    m = (rng.random(12000) + 0.5).astype(np.float64)
    M = meshoid_amd.Meshoid(pos, m, boxsize=box)
    O = oracle.OracleMeshoid(pos.astype(np.float64), m, boxsize=box)
    ngb = M.NearestNeighbors()
    dist = M.NeighborDistance()
    assert dist.dtype == np.float64 and ngb.dtype == np.int64
    # rows ascending, self first at distance 0 (cKDTree.query, Meshoid.py:157)
    assert (np.diff(dist, axis=1) >= 0).all()
    assert (ngb[:, 0] == np.arange(12000)).all() and (dist[:, 0] == 0).all()
    # the neighbour sets of the float64 kNN; distances are the float64 ones
    same = np.array([set(a) == set(b) for a, b in zip(ngb.tolist(), O.ngb.tolist())])
    assert same.all()
    np.testing.assert_allclose(dist, O.ngbdist, rtol=1e-15, atol=0)
    # h is HsmlIter's float64 root (refined over the final neighbours), so
    # h, density and volume agree with the oracle's to float64 rounding
    np.testing.assert_allclose(M.SmoothingLength(), O.kernel_radius, rtol=1e-12)
    np.testing.assert_allclose(M.Density(), O.density, rtol=1e-11)
    np.testing.assert_allclose(M.Volume(), O.vol, rtol=1e-11)


@pytest.mark.parametrize("order", [1, 2])
def test_compute_dweights_split(oracle, order):
    _torch()
    import meshoid_amd

    say("ComputeDWeights order", order)
    rng = np.random.default_rng(12)
    pos = rng.random((6000, 3))
    M = meshoid_amd.Meshoid(pos)
    M.ComputeDWeights(order)
    ngb, h = M.ngb, M.kernel_radius
    # the oracle's gradient_weights on the same rows (same columns, same h)
    G_o = oracle.gradient_weights(pos, ngb, h, np.arange(6000), order=order)
    W = oracle.kernel(M.ngbdist / h[:, None])
    nd = 3 if order == 1 else 9
    dx = pos[ngb] - pos[:, None, :]
    A = dx if order == 1 else np.concatenate([dx, 0.5 * dx**2, dx * np.roll(dx, -1, axis=2)], axis=2)
    cond = np.linalg.cond(np.einsum("ij,ijk,ijl->ikl", W, A, A))
    if order == 1:
        G = M.dweights
        assert M.d2weights is None and M.dweights_3rdorder is None
        assert G.shape == (6000, 3, 32) and G.dtype == np.float64
    else:
        assert M.d2weights.shape == (6000, 6, 32) and M.dweights_3rdorder.shape == (6000, 3, 32)
        # the split of Meshoid.py:134-140: rows dim.. are second derivatives
        G = np.concatenate([M.dweights_3rdorder, M.d2weights], axis=1)
    ok = np.isfinite(G).all(axis=(1, 2))
    assert ok.mean() > 0.99
    err = np.abs(G - G_o).reshape(6000, -1).max(1) / np.abs(G_o).reshape(6000, -1).max(1)
    say("max weight error / bound", np.max(err[ok] / np.maximum(1e-10, 1e-14 * cond[ok])))
    assert (err[ok] <= np.maximum(1e-10, 1e-14 * cond[ok])).all()
    assert G.shape[1] == nd


def _two_shell_stencil(a, b, c=None):
    pts = [np.zeros(3)]
    for r in (a, b):
        for ax in range(3):
            for sgn in (1.0, -1.0):
                e = np.zeros(3)
                e[ax] = sgn * r
                pts.append(e)
    if c is not None:
        for ax in range(3):
            for s1 in (1.0, -1.0):
                for s2 in (1.0, -1.0):
                    e = np.zeros(3)
                    e[ax], e[(ax + 1) % 3] = s1 * c, s2 * c
                    pts.append(e)
    return np.array(pts)


@pytest.mark.parametrize("h", [1.0, 0.9])
def test_gradient_weights_kernel_weighting_closed_form(oracle, h):
    """D(x^3)_x = (w_a a^4 + w_b b^4) / (w_a a^2 + w_b b^2) with w = W(r/h) on
    the symmetric two-shell stencil (order 1), and the order-2 analogue with
    the 12 edge points at |r| = sqrt(2) c."""
    _torch()
    import meshoid_amd

    a, b, c = 0.3, 0.8, 0.45
    for order, pos in ((1, _two_shell_stencil(a, b)), (2, _two_shell_stencil(a, b, c))):
        k = len(pos)
        ngb = np.arange(k, dtype=np.int32)[None, :]
        G = meshoid_amd.gradient_weights(pos, ngb, np.array([h]), np.array([0]), order=order)
        D = np.einsum("kj,j->k", G[0], pos[:, 0] ** 3)
        wa, wb = oracle.kernel(a / h), oracle.kernel(b / h)
        if order == 1:
            want = (wa * a**4 + wb * b**4) / (wa * a**2 + wb * b**2)
        else:
            wc = oracle.kernel(np.sqrt(2) * c / h)
            want = (2 * wa * a**4 + 2 * wb * b**4 + 8 * wc * c**4) / (2 * wa * a**2 + 2 * wb * b**2 + 8 * wc * c**2)
        say("order", order, "D(x^3)_x", D[0], "closed form", want)
        np.testing.assert_allclose(D[0], want, rtol=1e-12)
        assert np.abs(D[1:]).max() < 1e-12
