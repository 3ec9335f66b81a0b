"""CPU checks of the oracle's radtransfer restatement
(src/meshoid/radiation/radtransfer.py:79-177) against closed forms: pure
absorption attenuates the background by exp(-sum tau), the optically thin
limit adds j m W / (4 pi h^2), and a near absorber dims a far emitter (the
sweep runs farthest first, largest z first)."""
import numpy as np

from oracle import oracle as orc

C2 = 40 / (7 * np.pi)


def test_pure_absorber_on_background():
    # one particle on pixel (8, 8) of a 17 x 17 grid, L = 16 (dx = 1)
    img = orc.radtransfer(np.zeros((1, 1)), [2.0], [[0.5]], [[0.0, 0.0, 3.0]], [2.0], 17, 16.0, 0, i0=3.0)
    tau0 = 0.5 * C2 * 2.0 / 4.0
    assert np.isclose(img[8, 8, 0], 3.0 * np.exp(-tau0), rtol=1e-14)
    assert img[0, 0, 0] == 3.0  # outside the footprint


def test_thin_limit_and_bands():
    rng = np.random.default_rng(0)
    n = 50
    x = rng.random((n, 3)) * 6 + 5
    m = rng.random(n) + 0.5
    h = rng.random(n) + 1.0
    j = rng.random((n, 2))
    kap = np.full((n, 2), 1e-9)
    img = orc.radtransfer(j, m, kap, x, h, 33, 16.0, center=[8.0, 8.0, 0.0])
    dx = 0.5
    g = np.arange(33) * dx - 8.0 + 8.0  # pixel coordinates
    want = np.zeros((33, 33, 2))
    for i in range(n):
        r = np.sqrt((g[:, None] - x[i, 0]) ** 2 + (g[None, :] - x[i, 1]) ** 2)
        w = np.where(r <= h[i], orc.kernel2d(np.minimum(r / h[i], 1)), 0)
        want += (w * m[i] / h[i] ** 2)[:, :, None] * j[i][None, None, :] / (4 * np.pi)
    assert np.allclose(img, want, rtol=1e-6, atol=1e-12 * want.max())


def test_order_far_first():
    # emitter at z = 1 (far), absorber at z = 0 (near): the absorber dims it
    x = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 0.0]])
    j = np.array([[4.0], [0.0]])
    kap = np.array([[1.0], [2.0]])
    img = orc.radtransfer(j, [1.0, 1.0], kap, x, [1.0, 1.0], 3, 2.0, 0)
    tau_e, tau_a = C2, 2 * C2
    emit = -np.expm1(-tau_e) * 4.0 / (4 * np.pi * 1.0 + 1e-100)
    assert np.isclose(img[1, 1, 0], np.exp(-tau_a) * emit, rtol=1e-14)
    # swapping the depths puts the emitter in front: undimmed
    x2 = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    img2 = orc.radtransfer(j, [1.0, 1.0], kap, x2, [1.0, 1.0], 3, 2.0, 0)
    assert np.isclose(img2[1, 1, 0], emit, rtol=1e-14)
