"""CPU checks of the oracle's kernel-statistics restatements
(get_kernel_weights / KernelAverage / KernelVariance / Integrate,
src/meshoid/Meshoid.py:172-177,333-381) against closed forms and explicit
loops."""
import numpy as np

from oracle import oracle as orc


def _lists(n=200, k=8, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    d, ngb = orc.knn(pos, k)
    h = orc.hsml_iter(d, error_norm=1e-13)
    return pos, d, ngb, h


def test_kernel_weights_rows_and_values():
    pos, d, ngb, h = _lists()
    w = orc.kernel_weights(d, h)
    assert np.allclose(w.sum(1), 1.0, rtol=0, atol=1e-14)
    # self (q = 0) carries W(0) = 1 before the normalisation
    wsum = sum(orc.kernel(d[:, j] / h) for j in range(d.shape[1]))
    assert np.allclose(w[:, 0], 1.0 / wsum, rtol=1e-14)
    # one row by explicit cubic-spline arithmetic
    q = d[5] / h[5]
    W = np.array([1 - 6 * x * x + 6 * x**3 if x <= 0.5 else (2 * (1 - x) ** 3 if x <= 1 else 0.0) for x in q])
    assert np.allclose(w[5], W / W.sum(), rtol=1e-14)


def test_kernel_average_invariants():
    pos, d, ngb, h = _lists(seed=1)
    c = np.full(len(pos), 3.25)
    assert np.allclose(orc.kernel_average(d, h, ngb, c), 3.25, rtol=1e-14)
    f = np.random.default_rng(2).random(len(pos))
    i = 17
    w = orc.kernel(d[i] / h[i])
    expect = sum(w[j] * f[ngb[i, j]] for j in range(ngb.shape[1])) / w.sum()
    assert np.isclose(orc.kernel_average(d, h, ngb, f)[i], expect, rtol=1e-13)


def test_kernel_variance_explicit():
    pos, d, ngb, h = _lists(seed=3)
    f = np.random.default_rng(4).random((len(pos), 2))
    sd = orc.kernel_variance(ngb, f)
    i = 9
    vals = f[ngb[i]]
    mu = vals.sum(0) / len(vals)
    assert np.allclose(sd[i], np.sqrt(((vals - mu) ** 2).sum(0) / len(vals)), rtol=1e-13)
    assert np.allclose(orc.kernel_variance(ngb, np.ones(len(pos))), 0.0, atol=0)


def test_integrate_closed_forms():
    vol = np.random.default_rng(5).random(50)
    assert np.isclose(orc.integrate(vol, np.ones(50)), vol.sum(), rtol=1e-15)
    f = np.random.default_rng(6).random((50, 3))
    assert np.allclose(orc.integrate(vol, f), sum(vol[i] * f[i] for i in range(50)), rtol=1e-13)
