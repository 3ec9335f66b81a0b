"""GPU parity of the derivative operators and the Meshoid drop-in against the
FP64 oracle (same inputs, same neighbour sets)."""
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


def rel_err(a, b):
    """max |a-b| over rows, relative to the row's |b| (vector norm)."""
    a = np.asarray(a, dtype=np.float64).reshape(len(a), -1)
    b = np.asarray(b, dtype=np.float64).reshape(len(b), -1)
    scale = np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    return np.max(np.linalg.norm(a - b, axis=1) / scale)


@pytest.mark.parametrize("dtype,box", [(np.float64, None), (np.float32, None), (np.float32, 1.0),
                                       (np.float64, 1.0)])
def test_meshoid_D_matches_oracle(oracle, dtype, box):
    _torch()
    import meshoid_amd
    from meshoid_amd import synthetic

    say("D", dtype.__name__, box)
    pos = synthetic.uniform(20000, seed=1, dtype=dtype)
    f = np.sin(2 * np.pi * pos[:, 0]) * np.cos(2 * np.pi * pos[:, 1]) + pos[:, 2] ** 2
    M = meshoid_amd.Meshoid(pos, boxsize=box)
    O = oracle.OracleMeshoid(pos.astype(np.float64), boxsize=box)
    g = M.D(f)
    g_o = O.D(f.astype(np.float64))
    say("rel err D", rel_err(g, g_o))
    assert rel_err(g, g_o) <= 1e-5
    np.testing.assert_allclose(M.kernel_radius, O.kernel_radius, rtol=1e-6)
    np.testing.assert_allclose(M.density, O.density, rtol=4e-6)
    assert (M.ngb[:, 0] == np.arange(20000)).all()


def test_meshoid_D2_and_order2_match_oracle(oracle):
    _torch()
    import meshoid_amd

    rng = np.random.default_rng(42)
    x = rng.random((20000, 3))
    M = meshoid_amd.Meshoid(x, des_ngb=100)
    O = oracle.OracleMeshoid(x, des_ngb=100)
    comp = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 2): 4, (0, 2): 5}
    for (i, j), c in comp.items():
        f = x[:, i] * x[:, j]
        d2 = M.D2(f)
        expect = np.zeros(6)
        expect[c] = 2.0 if i == j else 1.0
        # reference known answer (test/test_gradients.ipynb): exact on quadratics
        assert np.abs(d2 - expect).max() < 2e-5, (i, j, np.abs(d2 - expect).max())
        assert rel_err(M.D(f, order=2), O.D(f, order=2)) <= 1e-5
    f = np.sin(3 * x[:, 0]) * np.exp(x[:, 1]) + x[:, 2] ** 3
    assert rel_err(M.D2(f), O.D2(f)) <= 1e-5


def test_vector_fields_curl_div(oracle):
    _torch()
    import meshoid_amd

    rng = np.random.default_rng(5)
    pos = rng.random((20000, 3))
    Om = np.array([0.5, -1.0, 2.0])
    v = np.cross(Om, pos - 0.5) + 0.01 * rng.standard_normal((20000, 3))
    M = meshoid_amd.Meshoid(pos)
    O = oracle.OracleMeshoid(pos)
    assert rel_err(M.D(v), O.D(v)) <= 1e-5
    np.testing.assert_allclose(M.Curl(v), O.Curl(v), atol=2e-5 * np.abs(O.Curl(v)).max())
    np.testing.assert_allclose(M.Div(v), O.Div(v), atol=2e-5 * np.abs(O.Curl(v)).max())
    c = M.Curl(np.cross(Om, pos - 0.5))
    # bug-compatible sign (Meshoid.py:311-315): (-2Wx, 2Wy, -2Wz)
    assert np.abs(c - [-2 * Om[0], 2 * Om[1], -2 * Om[2]]).max() < 1e-4
    c2 = M.Curl(np.cross(Om, pos - 0.5), correct_sign=True)
    assert np.abs(c2 - 2 * Om).max() < 1e-4


def test_masked_meshoid_and_quirks(oracle):
    _torch()
    import meshoid_amd

    rng = np.random.default_rng(8)
    pos = rng.random((10000, 3))
    # Meshoid.py:169 indexes the already-masked masses with the mask again, so
    # the reference only accepts masks whose indices are < Nmask
    with pytest.raises(IndexError):
        meshoid_amd.Meshoid(pos, particle_mask=np.arange(0, 10000, 3))
    with pytest.raises(IndexError):
        oracle.OracleMeshoid(pos, particle_mask=np.arange(0, 10000, 3))
    mask = rng.permutation(5000)
    m = rng.random(10000) + 0.5
    M = meshoid_amd.Meshoid(pos, m, particle_mask=mask, des_ngb=32)
    O = oracle.OracleMeshoid(pos, m, particle_mask=mask, des_ngb=32)
    np.testing.assert_allclose(M.density, O.density, rtol=4e-6)
    f = pos[:, 0] ** 2 + pos[:, 1]
    assert rel_err(M.D(f), O.D(f)) <= 1e-5
    with pytest.raises(IndexError):
        M.Div(rng.random((10000, 3)))  # Meshoid.py:330 indexes v[mask] with global ngb
    bm = np.zeros(10000, dtype=bool)
    bm[:4000] = True
    Mb = meshoid_amd.Meshoid(pos, particle_mask=bm)
    Ob = oracle.OracleMeshoid(pos, particle_mask=bm)
    assert rel_err(Mb.D(f), Ob.D(f)) <= 1e-5


def test_unweighted_and_multifield(oracle):
    _torch()
    import meshoid_amd

    rng = np.random.default_rng(2)
    pos = rng.random((15000, 3)).astype(np.float32)
    F = rng.standard_normal((15000, 2, 3)).astype(np.float32) + pos[:, None, :1]
    M = meshoid_amd.Meshoid(pos)
    O = oracle.OracleMeshoid(pos.astype(np.float64))
    d = M.D(F, weighted=False)
    d_o = O.D(F.astype(np.float64), weighted=False)
    assert d.shape == (15000, 2, 3, 3)
    assert rel_err(d, d_o) <= 1e-5


def test_gradient_weights_and_hsmliter_functions(oracle):
    torch = _torch()
    import meshoid_amd

    rng = np.random.default_rng(3)
    pos = rng.random((5000, 3))
    d_o, n_o = oracle.knn(pos, 32)
    h_o = oracle.hsml_iter(d_o, error_norm=1e-13)
    h = meshoid_amd.HsmlIter(d_o, error_norm=1e-13)
    np.testing.assert_allclose(h, h_o, rtol=1e-12)
    W = oracle.kernel(d_o / h_o[:, None])
    for order in (1, 2):
        G = meshoid_amd.gradient_weights(pos, n_o.astype(np.int32), h_o, np.arange(5000),
                                         order=order)
        G_o = oracle.gradient_weights(pos, n_o, h_o, np.arange(5000), order=order)
        bad = np.isnan(G).any(axis=(1, 2))
        # NaN only where the weighted fit is rank-deficient: fewer neighbours
        # with nonzero weight (self excluded) than unknowns
        nd = 3 if order == 1 else 9
        assert ((W[bad, 1:] > 0).sum(1) < nd).all()
        say("order", order, "rank-deficient rows", int(bad.sum()))
        # per-row bound: both solves are backward stable, so they agree to
        # ~cond(E) * eps (LU with pivoting in the oracle, LDL^T on the GPU)
        dx = pos[n_o] - pos[:, None, :]
        A = dx if order == 1 else np.concatenate(
            [dx, 0.5 * dx**2, dx * np.roll(dx, -1, axis=2)], axis=2)
        E = np.einsum("ij,ijk,ijl->ikl", W, A, A)
        cond = np.linalg.cond(E)
        ok = ~bad
        err = np.abs(G - G_o).reshape(5000, -1).max(1) / np.abs(G_o).reshape(5000, -1).max(1)
        assert (err[ok] <= np.maximum(1e-10, 1e-14 * cond[ok])).all()
    q = np.linspace(0, 1.5, 31)
    np.testing.assert_allclose(meshoid_amd.Kernel(q), oracle.kernel(q), rtol=1e-15, atol=1e-16)
    assert isinstance(meshoid_amd.Kernel(torch.tensor(q, device="cuda")), torch.Tensor)


def test_singular_moment_matrix_raises():
    _torch()
    import meshoid_amd

    pos = np.zeros((64, 3))
    pos[:, 0] = np.arange(64.0)  # collinear: rank-1 moment matrices in 3D
    M = meshoid_amd.Meshoid(pos, des_ngb=8)
    with pytest.raises(np.linalg.LinAlgError):
        M.D(pos[:, 0])


def test_linear_exactness_clustered_large():
    """2^21 clustered float32 particles: D of a linear field is exact to
    float32 rounding (size-independent property)."""
    torch = _torch()
    from meshoid_amd import engine, synthetic

    pos = synthetic.clustered(1 << 21, seed=3, nclump=128, background=0.1)
    p = torch.from_numpy(pos).cuda()
    ix = engine.SpatialIndex().build(p)
    out = engine.knn_hsml(ix, 32, keep_dist=False)
    b = torch.tensor([0.3, -1.2, 2.5], device="cuda", dtype=torch.float64)
    f = ((p.double() - 50.0) @ b).contiguous()  # float64 field: exact differences
    d1, _, st = engine.apply_D(out["nlist"], f)
    err = (d1[:, 0, :].double() - b).norm(dim=1) / b.norm()
    say("linear exactness max rel err", float(err.max()), "median", float(err.median()))
    assert float(torch.quantile(err[:1 << 20], 0.999)) < 1e-5


@pytest.mark.parametrize("k,box,weighted,masked", [(32, None, True, False), (33, None, True, False),
                                                   (2, None, True, False), (32, 10.0, True, False),
                                                   (32, None, False, False), (32, None, True, True)])
def test_single_field_path_equals_generic(k, box, weighted, masked):
    """One float32 field in 3D (order 1) takes the pipelined packed-load
    kernel; the same field as column 0 of a two-field input takes the generic
    kernel.  Both sum the same float64 terms, split over 4 and 2 lanes per
    query: equal status and NaN rows, and float32 results that differ only
    where the float64 sums' last bits cross a float32 rounding boundary."""
    torch = _torch()
    from meshoid_amd import engine, synthetic

    say("single-field path", k, box, weighted, masked)
    n = 150000
    pos = synthetic.clustered(n, seed=k, nclump=16, box=10.0, background=0.1, dtype=np.float64)
    if box:
        box = 10.0
        pos = np.mod(pos, box)
    pos = pos.astype(np.float32)
    if box:
        pos[pos >= box] = 0.0
    rng = np.random.default_rng(k)
    f = rng.standard_normal(n).astype(np.float32)
    ix = engine.SpatialIndex().build(torch.from_numpy(pos).cuda(), boxsize=box)
    qidx = None
    if masked:
        qidx = torch.from_numpy(np.sort(rng.choice(n, n // 3, replace=False))).cuda()
    nl = engine.knn_hsml(ix, k, qidx=qidx)["nlist"]
    fd = torch.from_numpy(f).cuda()
    d1, _, st1 = engine.apply_D(nl, fd[:, None].contiguous(), weighted=weighted)
    d2, _, st2 = engine.apply_D(nl, torch.stack([fd, 2 * fd], 1).contiguous(), weighted=weighted)
    torch.cuda.synchronize()
    a = d1[:, 0].cpu().numpy()
    b = d2[:, 0].cpu().numpy()
    assert torch.equal(st1, st2)
    nan = np.isnan(a).any(1)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    same = (a[~nan] == b[~nan]).mean()
    err = rel_err(a[~nan], b[~nan]) if (~nan).any() else 0.0
    say("bit-identical fraction", same, "max row rel diff", err)
    assert same > 0.999 and err < 1e-5
