"""Pins the FP64 oracle against the reference's known answers, scipy golden
vectors and the mathematics (closed forms, invariants, brute force)."""
import glob
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _same_sets_up_to_ties(ngb_a, d_a, ngb_b, d_b, rtol=2.3e-16):
    """Rows hold the same neighbour set, except members whose distance ties the
    k-th distance to within 1 ulp."""
    bad = 0
    for i in range(ngb_a.shape[0]):
        sa, sb = set(ngb_a[i].tolist()), set(ngb_b[i].tolist())
        if sa == sb:
            continue
        kth = max(d_a[i, -1], d_b[i, -1])
        diff = sa ^ sb
        da = dict(zip(ngb_a[i].tolist(), d_a[i].tolist()))
        db = dict(zip(ngb_b[i].tolist(), d_b[i].tolist()))
        for j in diff:
            dj = da.get(j, db.get(j))
            if abs(dj - kth) > 2 * rtol * kth:
                bad += 1
    return bad


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "knn_*.npz"))))
def test_oracle_knn_matches_ckdtree_golden(oracle, path):
    z = np.load(path)
    pos = z["pos"].astype(np.float64)
    k = int(z["k"])
    box = float(z["boxsize"])
    q = pos[z["qidx"]]
    d, n = oracle.knn(pos, k, box if box > 0 else None, queries=q)
    assert _same_sets_up_to_ties(n, d, z["ngb"], z["dist"]) == 0
    np.testing.assert_allclose(d, z["dist"], rtol=1e-15, atol=0)
    assert (n[:, 0] == z["qidx"]).all()  # self first, at distance 0


@pytest.mark.parametrize("dim,box", [(3, None), (3, 1.0), (2, None), (1, None), (2, 1.0)])
def test_oracle_knn_matches_brute_force(oracle, dim, box):
    rng = np.random.default_rng(dim * 10 + (box is not None))
    pos = rng.random((600, dim))
    d1, n1 = oracle.knn(pos, 16, box)
    d2, n2 = oracle.knn(pos, 16, box, brute=True)
    np.testing.assert_array_equal(d1, d2)
    assert _same_sets_up_to_ties(n1, d1, n2, d2) == 0


def test_oracle_knn_missing_neighbours(oracle):
    pos = np.random.default_rng(0).random((5, 3))
    d, n = oracle.knn(pos, 8)
    assert (n[:, 5:] == 5).all() and np.isinf(d[:, 5:]).all()


def _q_star(k, dim):
    """Closed form root for the row [0, r, ..., r]: norm*(1 + (k-1) W(q)) = k."""
    norm = {3: 32 / 3, 2: 40 / 7, 1: 8 / 3}[dim]
    Wt = (k / norm - 1) / (k - 1)
    if Wt <= 0.25:  # outer branch 2(1-q)^3
        return 1 - (Wt / 2) ** (1 / 3)
    from scipy.optimize import brentq
    return brentq(lambda q: 1 - 6 * q**2 + 6 * q**3 - Wt, 0, 0.5)


@pytest.mark.parametrize("k,dim", [(32, 3), (64, 3), (20, 2), (100, 3)])
def test_oracle_hsml_closed_form(oracle, k, dim):
    r = np.array([0.3, 1.0, 7.5])
    rows = np.zeros((3, k))
    rows[:, 1:] = r[:, None]
    h = oracle.hsml_iter(rows, dim=dim, error_norm=1e-13)
    np.testing.assert_allclose(h, r / _q_star(k, dim), rtol=1e-13)
    if (k, dim) == (32, 3):
        assert abs(_q_star(32, 3) - 0.6816686321542267) < 1e-15


def test_oracle_hsml_invariants(oracle):
    rng = np.random.default_rng(7)
    pos = rng.random((3000, 3))
    d, _ = oracle.knn(pos, 32)
    h, it = oracle.hsml_iter(d, error_norm=1e-13, return_iters=True)
    W = oracle.kernel(d / h[:, None])
    n = 32 / 3 * W.sum(1)
    assert np.max(np.abs(n - 32)) < 1e-12
    assert (h >= d[:, 1]).all() and (h <= 1.5638742363463305 * d[:, -1]).all()
    assert it.max() < 80
    h2 = oracle.hsml_iter(d * 3.0, error_norm=1e-13)
    np.testing.assert_allclose(h2, 3 * h, rtol=1e-12)


def test_oracle_kernel_values(oracle):
    q = np.array([0.0, 0.25, 0.5, 0.75, 1.0, 1.5])
    expect = [1.0, 1 - 6 / 16 + 6 / 64, 0.25, 2 / 64, 0.0, 0.0]
    np.testing.assert_allclose(oracle.kernel(q), expect, rtol=0, atol=1e-16)
    for x in q:
        assert oracle.lib().orc_kernel(float(x)) == oracle.kernel(x)


@pytest.mark.parametrize("box", [None, 1.0])
def test_oracle_gradient_linear_exactness(oracle, box):
    rng = np.random.default_rng(11)
    pos = rng.random((2000, 3))
    M = oracle.OracleMeshoid(pos, boxsize=box)
    G = oracle.gradient_weights(pos, M.ngb, M.kernel_radius, M.particle_mask, boxsize=box)
    dx = pos[M.ngb] - pos[:, None, :]
    if box:
        dx = np.where(np.abs(dx) > box / 2, -np.copysign(box - np.abs(dx), dx), dx)
    ident = np.einsum("ikj,ijl->ikl", G, dx)
    assert np.abs(ident - np.eye(3)).max() < 1e-10
    if box is None:
        b = np.array([0.3, -1.2, 2.5])
        f = 4.0 + pos @ b
        np.testing.assert_allclose(M.D(f), np.broadcast_to(b, (2000, 3)), atol=1e-10)


def test_oracle_D2_quadratic_exactness(oracle):
    """test/test_gradients.ipynb cell 0: D2(x_i x_j) is exactly 2 (i==j) or 1 on
    the matching component and ~0 elsewhere (order xx,yy,zz,xy,yz,zx)."""
    rng = np.random.default_rng(42)
    x = rng.random((3000, 3))
    M = oracle.OracleMeshoid(x, des_ngb=100)
    comp = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 2): 4, (0, 2): 5}
    for (i, j), c in comp.items():
        d2 = M.D2(x[:, i] * x[:, j])
        expect = np.zeros(6)
        expect[c] = 2.0 if i == j else 1.0
        assert np.abs(d2 - expect).max() < 1e-9
        d1 = M.D(x[:, i] * x[:, j], order=2)
        g = np.zeros((3000, 3))
        g[:, i] += x[:, j]
        g[:, j] += x[:, i]
        assert np.abs(d1 - g).max() < 1e-9


def test_oracle_density_and_curl_div(oracle):
    rng = np.random.default_rng(5)
    pos = rng.random((2000, 3))
    m = rng.random(2000) + 0.5
    M = oracle.OracleMeshoid(pos, m)
    np.testing.assert_allclose(M.density * 4 * np.pi / 3 * M.kernel_radius**3 / (32 * m), 1,
                               rtol=1e-14)
    Om = np.array([0.5, -1.0, 2.0])
    v = np.cross(Om, pos - 0.5)
    assert np.abs(M.Div(v)).max() < 1e-10
    np.testing.assert_allclose(M.Curl(v), np.broadcast_to([-2 * Om[0], 2 * Om[1], -2 * Om[2]],
                                                          (2000, 3)), atol=1e-10)


def test_oracle_singular_moment_matrix(oracle):
    # all neighbours on a plane: every offset has dz = 0 exactly, so the z row
    # and column of E are exactly zero and the elimination meets an exact zero
    # pivot on every row (np.linalg.inv raises LinAlgError there)
    rng = np.random.default_rng(1)
    pos = rng.random((200, 3))
    pos[:, 2] = 0.5
    d, n = oracle.knn(pos, 32)
    h = oracle.hsml_iter(d)
    G, st = oracle.gradient_weights(pos, n, h, np.arange(200), return_status=True)
    assert (st == 1).all() and np.isnan(G).all()
    # the same rows lifted off the plane are regular
    pos[:, 2] = 0.5 + 0.01 * rng.standard_normal(200)
    d, n = oracle.knn(pos, 32)
    G, st = oracle.gradient_weights(pos, n, oracle.hsml_iter(d), np.arange(200), return_status=True)
    assert (st == 0).all() and np.isfinite(G).all()


def _two_shell_stencil(a, b, c=None):
    """Query at the origin (row 0) with neighbours at +-a and +-b on every
    axis and, optionally, at (+-c, +-c, 0) and its cyclic images."""
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
def test_oracle_weighting_closed_form_order1(oracle, h):
    """gradient_weights weights each pair by W(|dx|/h) (derivatives.py:50-51).
    On the symmetric two-shell stencil E = diag(2 w_a a^2 + 2 w_b b^2) and
    D(x^3)_x = (w_a a^4 + w_b b^4) / (w_a a^2 + w_b b^2): an unweighted fit,
    W(r) or W(r/(2h)) each give a different value."""
    a, b = 0.3, 0.8
    pos = _two_shell_stencil(a, b)
    k = len(pos)
    ngb = np.arange(k)[None, :]
    G = oracle.gradient_weights(pos, ngb, np.array([h]), np.array([0]))
    f = pos[:, 0] ** 3
    D = oracle.apply_weights(G, ngb, np.array([0]), f)[0]
    wa, wb = oracle.kernel(a / h), oracle.kernel(b / h)
    want = (wa * a**4 + wb * b**4) / (wa * a**2 + wb * b**2)
    np.testing.assert_allclose(D, [want, 0.0, 0.0], rtol=1e-13, atol=1e-15)
    # the alternatives a dropped or mis-scaled weight would give
    for alt in (1.0, 1.0), (oracle.kernel(a / (2 * h)), oracle.kernel(b / (2 * h))):
        wrong = (alt[0] * a**4 + alt[1] * b**4) / (alt[0] * a**2 + alt[1] * b**2)
        assert abs(wrong - want) > 1e-3
    Gu = oracle.gradient_weights(pos, ngb, np.array([h]), np.array([0]), weighted=False)
    Du = oracle.apply_weights(Gu, ngb, np.array([0]), f)[0]
    np.testing.assert_allclose(Du[0], (a**4 + b**4) / (a**2 + b**2), rtol=1e-13)


def test_oracle_weighting_closed_form_order2(oracle):
    """Order 2 (derivatives.py:83-101) on the two-shell stencil plus the 12
    edge points (+-c, +-c, 0): the reflection symmetries split the 9x9 system
    by parity, so the x row is (2 w_a a^4 + 2 w_b b^4 + 8 w_c c^4) /
    (2 w_a a^2 + 2 w_b b^2 + 8 w_c c^2) with w_c = W(sqrt(2) c / h), and the
    second-derivative rows of the odd field x^3 vanish."""
    a, b, c, h = 0.3, 0.8, 0.45, 1.0
    pos = _two_shell_stencil(a, b, c)
    k = len(pos)
    ngb = np.arange(k)[None, :]
    G, st = oracle.gradient_weights(pos, ngb, np.array([h]), np.array([0]), order=2, return_status=True)
    assert st[0] == 0
    D = oracle.apply_weights(G, ngb, np.array([0]), pos[:, 0] ** 3)[0]
    wa, wb, wc = oracle.kernel(a / h), oracle.kernel(b / h), oracle.kernel(np.sqrt(2) * c / h)
    want = (2 * wa * a**4 + 2 * wb * b**4 + 8 * wc * c**4) / (2 * wa * a**2 + 2 * wb * b**2 + 8 * wc * c**2)
    np.testing.assert_allclose(D[0], want, rtol=1e-12)
    assert np.abs(D[1:]).max() < 1e-12
    # and the fit is exact on the quadratic part of the basis
    D2 = oracle.apply_weights(G, ngb, np.array([0]), pos[:, 0] * pos[:, 1])[0]
    np.testing.assert_allclose(D2, [0, 0, 0, 0, 0, 0, 1, 0, 0], atol=1e-12)


def test_oracle_curl_mask_quirk_raises(oracle):
    pos = np.random.default_rng(3).random((300, 3))
    M = oracle.OracleMeshoid(pos, particle_mask=np.arange(100))
    with pytest.raises(IndexError):
        M.Div(np.random.default_rng(0).random((300, 3)))
