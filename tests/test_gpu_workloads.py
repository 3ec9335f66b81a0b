"""The precision, batched and learned-seed workloads at their full sizes,
checked against the FP64 oracle on samples (not only against the single-set
GPU path):

* precision form: 2^24 clustered particles (256 clumps + 10% background),
  order-2 fits (Meshoid.D2 and D(order=2), src/meshoid/Meshoid.py:264-295) of
  f = sin(kx) sin(ky) sin(kz), k = 2 pi / 10, and of a Gaussian random field
  (P(k) ~ k^-3 on a periodic 256^3 grid), on 65 536 sampled rows; the bound is
  stated for rows whose equilibrated 9x9 moment matrix has cond <= 1e8 (there a
  float64 solve is meaningful; beyond it both solves return noise);
* batched form: 64 snapshots, N log-uniform in [2^14, 2^20], K = 32 fields,
  des_ngb alternating 32 / 64, through one graph launch; 4096 rows of the
  smallest and of the largest snapshot against the oracle;
* learned-seed form: 2^24 clustered particles, des_ngb = 32: the seeded search
  returns the plain search's neighbour sets and smoothing lengths.
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


def equilibrated_cond(p64, ngb, idx, h):
    """cond of each row's weighted 9x9 moment matrix (derivatives.py:83-101)
    after symmetric diagonal equilibration."""
    from oracle import oracle as orc

    dx = p64[ngb] - p64[idx][:, None, :]
    w = orc.kernel(np.linalg.norm(dx, axis=2) / h[:, None])
    x, y, z = dx[..., 0], dx[..., 1], dx[..., 2]
    a = np.stack([x, y, z, x * x / 2, y * y / 2, z * z / 2, x * y, y * z, z * x], axis=2)
    E = np.einsum("ij,ija,ijb->iab", w, a, a)
    dg = np.sqrt(np.einsum("iaa->ia", E))
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.linalg.cond(E / (dg[:, :, None] * dg[:, None, :]))
    return np.where(np.isfinite(c), c, np.inf)


def test_precision_form_2e24_order2(oracle):
    torch = _torch()
    from meshoid_amd import engine, synthetic

    n = 1 << 24
    say("precision form: generating 2^24 clustered particles + GRF")
    pos = synthetic.clustered(n, seed=0, nclump=256, background=0.1)
    f, _ = synthetic.sin3(pos)
    grf = synthetic.grf(pos, seed=7)
    F = np.stack([f, grf], axis=1).astype(np.float32)
    pos_d = torch.from_numpy(pos).cuda()
    ix = engine.SpatialIndex().build(pos_d)
    out = engine.knn_hsml(ix, 32, keep_dist=False)
    say("order-2 fits on the GPU")
    o1, o2, st = engine.apply_D(out["nlist"], torch.from_numpy(F).cuda(), order=2, want_d2=True)
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(n, 1 << 16, replace=False))
    o1 = o1.cpu().numpy()[idx]
    o2 = o2.cpu().numpy()[idx]
    h = out["h"].cpu().numpy()[idx]
    say("oracle on 65536 rows")
    p64 = pos.astype(np.float64)
    d, ngb = oracle.knn(p64, 32, queries=p64[idx])
    h_o = oracle.hsml_iter(d, error_norm=1e-13)
    assert np.abs(h / h_o - 1).max() <= 1e-6
    G = oracle.gradient_weights(p64, ngb, h_o, idx, order=2)
    ref = oracle.apply_weights(G, ngb, idx, F.astype(np.float64))  # (m, 2, 9)
    cond = equilibrated_cond(p64, ngb, idx, h_o)
    ok = cond <= 1e8
    say("rows with cond <= 1e8:", int(ok.sum()), "of", len(idx))
    assert ok.mean() > 0.95
    for j, name in enumerate(("sin", "grf")):
        eg = rel_rows(o1[:, j], ref[:, j, :3])
        eh = rel_rows(o2[:, j], ref[:, j, 3:])
        say(name, "order-2 grad max rel", eg[ok].max(), "median", np.median(eg[ok]),
            "| hessian max rel", eh[ok].max(), "median", np.median(eh[ok]))
        assert eg[ok].max() <= 1e-5, name
        assert eh[ok].max() <= 1e-5, name


def test_batched_form_64_snapshots_vs_oracle(oracle):
    torch = _torch()
    from meshoid_amd import synthetic
    from meshoid_amd.batch import SnapshotBatch

    say("batched form: generating 64 snapshots, K=32")
    snaps = synthetic.snapshot_batch(64, K=32, seed=0)
    sizes = [p.shape[0] for p, _, _ in snaps]
    ks = [k for _, _, k in snaps]
    say("sizes", min(sizes), "...", max(sizes), "total", sum(sizes))
    B = SnapshotBatch(sizes, ks, 32, nbranch=4)
    pos_b = [torch.from_numpy(p).cuda() for p, _, _ in snaps]
    f_b = [torch.from_numpy(f).cuda() for _, f, _ in snaps]
    B.run(pos_b, f_b, graph=False)
    out = B.run(pos_b, f_b, graph=True)  # captured and replayed
    torch.cuda.synchronize()
    for s in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        p64 = snaps[s][0].astype(np.float64)
        k = ks[s]
        rng = np.random.default_rng(s)
        idx = np.sort(rng.choice(len(p64), min(4096, len(p64)), replace=False))
        say(f"snapshot {s}: n={len(p64)} k={k}, oracle on {len(idx)} rows")
        d, ngb = oracle.knn(p64, k, queries=p64[idx])
        h_o = oracle.hsml_iter(d, error_norm=1e-13)
        h = out["h"][s].cpu().numpy()[idx]
        assert np.abs(h / h_o - 1).max() <= 1e-6
        rho_o = k / (4 * np.pi / 3 * h_o**3)
        np.testing.assert_allclose(out["rho"][s].cpu().numpy()[idx], rho_o, rtol=4e-6)
        G = oracle.gradient_weights(p64, ngb, h_o, idx)
        ref = oracle.apply_weights(G, ngb, idx, snaps[s][1].astype(np.float64))  # (m, 32, 3)
        g = out["grad"][s].cpu().numpy()[idx]
        e = np.linalg.norm(g - ref, axis=2) / np.maximum(np.linalg.norm(ref, axis=2), 1e-300)
        say("  grad max rel", e.max(), "median", np.median(e))
        assert e.max() <= 1e-5


def test_learned_seed_2e24_equals_plain():
    torch = _torch()
    from meshoid_amd import engine, synthetic
    from meshoid_amd import seed as seedmod

    n = 1 << 24
    say("learned form: generating 2^24 clustered particles")
    pos = synthetic.clustered(n, seed=0, nclump=256, background=0.1)
    pos_d = torch.from_numpy(pos).cuda()
    ix = engine.SpatialIndex().build(pos_d)
    plain = engine.knn_hsml(ix, 32, keep_dist=False)
    model = seedmod.load_default(32)
    hp = model.predict(ix)
    say("seeded search")
    seeded = engine.knn_hsml(ix, 32, keep_dist=False, seed_h=hp, seed_r=(1.3 * hp).contiguous())
    dh = (seeded["h"] / plain["h"] - 1).abs().max().item()
    say("max rel h difference", dh)
    assert dh <= 1e-6
    a, _ = plain["nlist"].export(want_dist=False, f64=True)
    b, _ = seeded["nlist"].export(want_dist=False, f64=True)
    a = torch.sort(a, dim=1).values
    b = torch.sort(b, dim=1).values
    diff = int((a != b).any(dim=1).sum())
    say("rows with different neighbour sets", diff)
    assert diff == 0
    miss = int(((seeded["status"] & 128) != 0).sum())
    say("radius misses recomputed", miss)
