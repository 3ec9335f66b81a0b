"""Parity cases shaped like the reference's measurement configurations:

* 65 536 uniform float64 particles, f = sin(2 pi x) cos(2 pi y) + z^2: the GPU
  gradient equals the FP64 oracle's, and both approach the analytic gradient
  away from the faces (the reference's own walkthrough check, README.md:190-220);
* 4 194 304 clustered float32 particles (256 clumps), K = 4 fields (rho and a
  rotation field v = Omega x (r - 50) + N(0, 0.1^2)): h, rho and D of all four
  fields against the FP64 oracle on a random subset, and the rotation field's
  Div / Curl known answers (Div = 0, reference Curl = (-2 Ox, 2 Oy, -2 Oz)).
"""
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


def test_uniform_65536_analytic_gradient(oracle):
    _torch()
    import meshoid_amd

    say("config: 65536 uniform f64")
    rng = np.random.default_rng(0)
    pos = rng.random((65536, 3))
    x, y, z = pos.T
    f = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + z**2
    g_exact = np.stack([2 * np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y),
                        -2 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y), 2 * z], axis=1)
    M = meshoid_amd.Meshoid(pos, np.full(65536, 1.0 / 65536), des_ngb=32)
    g = M.D(f)
    O = oracle.OracleMeshoid(pos, np.full(65536, 1.0 / 65536), des_ngb=32)
    g_o = O.D(f)
    assert rel_rows(g, g_o).max() <= 1e-5
    h = M.SmoothingLength()
    inner = np.all((pos > 2 * h[:, None]) & (pos < 1 - 2 * h[:, None]), axis=1)
    e_gpu = rel_rows(g[inner], g_exact[inner])
    e_orc = rel_rows(g_o[inner], g_exact[inner])
    say("interior discretisation error median", np.median(e_gpu), "oracle", np.median(e_orc))
    # the same discrete operator: identical discretisation error to 1e-6
    assert abs(np.median(e_gpu) - np.median(e_orc)) <= 1e-6
    assert np.median(e_gpu) < 0.05


def test_clustered_4M_four_fields(oracle):
    torch = _torch()
    import meshoid_amd
    from meshoid_amd import engine, synthetic

    n = 4194304
    say("config: 4M clustered f32, K=4")
    pos = synthetic.clustered(n, seed=2, nclump=256, background=0.0)
    m = np.full(n, 1.0 / n)
    pos_d = torch.from_numpy(pos).cuda()
    ix = engine.SpatialIndex()
    ix.build(pos_d)
    out = engine.knn_hsml(ix, 32, m_rows=torch.from_numpy(m).cuda(), keep_dist=False)
    rho = out["rho"].cpu().numpy().astype(np.float64)
    v = synthetic.rotation_field(pos, seed=3)
    F = np.concatenate([rho[:, None], v], axis=1).astype(np.float32)
    d1, _, st = engine.apply_D(out["nlist"], torch.from_numpy(F).cuda())
    d1 = d1.cpu().numpy()
    h = out["h"].cpu().numpy()
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(n, 1 << 16, replace=False))
    p64 = pos.astype(np.float64)
    say("oracle on 65536 rows")
    d, ngb = oracle.knn(p64, 32, queries=p64[idx])
    h_o = oracle.hsml_iter(d, error_norm=1e-13)
    np.testing.assert_allclose(h[idx], h_o, rtol=1e-6)
    rho_o = 32 * m[idx] / (4 * np.pi / 3 * h_o**3)
    np.testing.assert_allclose(rho[idx], rho_o, rtol=4e-6)
    G = oracle.gradient_weights(p64, ngb, h_o, idx)
    g_o = oracle.apply_weights(G, ngb, idx, F.astype(np.float64))  # (m, 4, 3)
    for c in range(4):
        e = rel_rows(d1[idx, c], g_o[:, c])
        say("field", c, "max rel", e.max())
        assert e.max() <= 1e-5, c
    # rotation field: Div = 0 and the reference Curl signs, up to the noise
    dv = d1[:, 1:, :]
    div = dv[:, 0, 0] + dv[:, 1, 1] + dv[:, 2, 2]
    curl = np.stack([dv[:, 1, 2] - dv[:, 2, 1], dv[:, 0, 2] - dv[:, 2, 0], dv[:, 0, 1] - dv[:, 1, 0]], 1)
    ok = (st.cpu().numpy() & 64) == 0
    say("median |div|", np.median(np.abs(div[ok])), "median curl", np.median(curl[ok], axis=0))
    # noise of 0.1 over neighbour spacings: the medians are what the exact field gives
    assert abs(np.median(div[ok])) < 0.2
    np.testing.assert_allclose(np.median(curl[ok], axis=0), [0.0, 0.0, -2.0], atol=0.2)
