"""CPU checks of the oracle's Reconstruct / Slice / InterpToGrid
restatement (src/meshoid/Meshoid.py:383-555) against brute force and closed
forms: order 0 is the brute-force nearest value, order 1 reproduces linear
fields and order 2 quadratic ones (the reference walkthrough's "Reconstruct
orders 1-2 exact on x"), the grids put the reconstructed coordinate where the
grid geometry says, and the mask side effect is the sorted set of nearest
particles."""
import numpy as np

from oracle import oracle as orc


def _set(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    return rng, rng.random((n, 3))


def test_order0_is_bruteforce_nearest():
    rng, pos = _set(1500, 1)
    f = rng.standard_normal((1500, 2))
    tp = rng.random((400, 3)) * 1.2 - 0.1  # some points outside the particles' box
    O = orc.OracleMeshoid(pos, des_ngb=16)
    r0 = O.Reconstruct(f, tp, 0)
    d2 = ((tp[:, None, :] - pos[None]) ** 2).sum(-1)
    assert np.array_equal(r0, f[d2.argmin(1)])
    # order 0 leaves the mask alone (the reference returns before it)
    assert np.array_equal(O.particle_mask, np.arange(1500))


def test_order1_linear_exact_and_mask():
    rng, pos = _set()
    b = np.array([0.3, -2.0, 0.7])
    f = 1.5 + pos @ b
    tp = rng.random((500, 3)) * 0.8 + 0.1
    O = orc.OracleMeshoid(pos, des_ngb=32)
    r1 = O.Reconstruct(f, tp, 1)
    assert np.abs(r1 - (1.5 + tp @ b)).max() < 1e-12
    d2 = ((tp[:, None, :] - pos[None]) ** 2).sum(-1)
    u = np.unique(d2.argmin(1))
    assert np.array_equal(O.particle_mask, u)
    assert O.kernel_radius.shape == (len(u),)


def test_order2_quadratic_exact():
    rng, pos = _set(4000, 2)
    Q = np.array([[1.0, 0.2, 0.0], [0.2, -1.0, 0.5], [0.0, 0.5, 2.0]])
    b = np.array([0.3, -2.0, 0.7])
    quad = lambda x: np.einsum("ni,ij,nj->n", x, Q, x) + x @ b
    tp = rng.random((600, 3)) * 0.8 + 0.1
    O = orc.OracleMeshoid(pos, des_ngb=32)
    r2 = O.Reconstruct(quad(pos), tp, 2)
    err = np.abs(r2 - quad(tp))
    # exact up to rows whose 9x9 moment matrix is singular at rounding level
    # (fewer than ~10 neighbours inside h): there both FP64 solves are noise
    assert np.quantile(err, 0.98) < 1e-10
    # order 1 on the same quadratic is only first-order accurate
    O1 = orc.OracleMeshoid(pos, des_ngb=32)
    assert np.median(np.abs(O1.Reconstruct(quad(pos), tp, 1) - quad(tp))) > 1e-5


def test_vector_field_and_float32_dtype():
    rng, pos = _set(2000, 3)
    A = rng.standard_normal((3, 2))
    f = (pos @ A).astype(np.float32)
    tp = rng.random((300, 3)) * 0.8 + 0.1
    O = orc.OracleMeshoid(pos, des_ngb=32)
    r = O.Reconstruct(f, tp, 1)
    assert r.dtype == np.float32 and r.shape == (300, 2)
    assert np.abs(r - tp @ A).max() < 1e-5  # float32 rounding of the result


def test_slice_and_grid_geometry():
    rng, pos = _set(3000, 4)
    O = orc.OracleMeshoid(pos, des_ngb=32)
    c = np.array([0.5, 0.45, 0.55])
    ax = np.linspace(-0.3, 0.3, 9)
    for plane, (iu, iv, iw) in {"z": (0, 1, 2), "x": (1, 2, 0), "y": (0, 2, 1)}.items():
        for comp in range(3):
            s = O.Slice(pos[:, comp].copy(), 0.6, c, res=9, order=1, plane=plane)
            assert s.shape == (9, 9)
            want = {iu: c[iu] + ax[:, None] + 0 * ax[None, :], iv: c[iv] + ax[None, :] + 0 * ax[:, None],
                    iw: np.full((9, 9), c[iw])}[comp]
            assert np.abs(s - want).max() < 1e-12, (plane, comp)
    e = np.linspace(-0.2, 0.2, 5)
    mid = 0.5 * (e[1:] + e[:-1])
    g = O.InterpToGrid(pos[:, 1].copy(), 0.4, c, res=4, order=1)
    assert g.shape == (4, 4, 4)
    assert np.abs(g - (c[1] + mid[None, :, None])).max() < 1e-12


def test_periodic_nearest_image():
    rng = np.random.default_rng(5)
    pos = rng.random((2000, 3)) * 10.0
    O = orc.OracleMeshoid(pos, des_ngb=32, boxsize=10.0)
    tp = rng.random((200, 3)) * 10.0
    tp[:20, 0] = rng.random(20) * 0.05  # near the x = 0 face: nearest may lie across it
    r0 = O.Reconstruct(np.arange(2000.0), tp, 0)
    d = np.abs(tp[:, None, :] - pos[None])
    d = np.minimum(d, 10.0 - d)
    assert np.array_equal(r0, (d**2).sum(-1).argmin(1).astype(float))
    # a linear field is not periodic: away from every face (neighbourhoods
    # that do not wrap) order 1 reproduces it
    inner = np.all((tp > 2) & (tp < 8), axis=1)
    f = pos[:, 1] * 2.0
    O2 = orc.OracleMeshoid(pos, des_ngb=32, boxsize=10.0)
    r1 = O2.Reconstruct(f, tp[inner], 1)
    assert np.abs(r1 - 2.0 * tp[inner, 1]).max() < 1e-9
