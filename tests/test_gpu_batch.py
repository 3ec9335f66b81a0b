"""Batched form: every snapshot of a SnapshotBatch equals the single-set path
(the same index and kNN kernels bit for bit; the derivative contraction up to
float64 summation order), eager and through graph replays (bit for bit
against the eager batch), and the single-set path itself matches the FP64
oracle (tests/test_gpu_ops.py)."""
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


def _snapshots(torch, sizes, K, seed=0):
    from meshoid_amd import synthetic

    rng = np.random.default_rng(seed)
    pos, fld = [], []
    for i, n in enumerate(sizes):
        p = synthetic.clustered(n, seed=seed + i, nclump=int(rng.integers(8, 33)), box=10.0, background=0.1)
        x = p.astype(np.float64)
        f = rng.standard_normal((n, K)) + x @ rng.standard_normal((3, K)) + \
            np.einsum("ni,ij,nj->n", x, rng.standard_normal((3, 3)), x)[:, None] * rng.standard_normal(K)
        pos.append(torch.from_numpy(p).cuda())
        fld.append(torch.from_numpy(f.astype(np.float32)).cuda())
    return pos, fld


def same(a, b):
    """bit-for-bit equal, NaN where the other is NaN (rank-deficient rows)"""
    import torch

    torch.testing.assert_close(a, b, rtol=0, atol=0, equal_nan=True)
    return True


def moment_cond(pos, ngb, h):
    """cond of each row's weighted 3x3 moment matrix sum_j W(r_j/h) d_j d_j^T
    (derivatives.py:69-78), from the exported neighbour lists."""
    from oracle import oracle as orc

    p = pos.astype(np.float64)
    d = p[ngb] - p[:, None, :]
    w = orc.kernel(np.linalg.norm(d, axis=2) / h.astype(np.float64)[:, None])
    E = np.einsum("ij,ija,ijb->iab", w, d, d)
    with np.errstate(all="ignore"):
        c = np.linalg.cond(E)
    return np.where(np.isfinite(c), c, np.inf)


def close(g_batch, g_single, s, cond):
    """Same operator, float64 sums in another order: rows agree to
    1e-5 + 1e-13 cond(E) relative (per field: the solve's sensitivity to the
    summation order grows with the conditioning), NaN rows (rank-deficient)
    in the same places."""
    g1 = g_batch.cpu().numpy().astype(np.float64)
    g2 = g_single.cpu().numpy().astype(np.float64)
    ok = np.isfinite(g2).all(axis=(1, 2))
    assert (np.isfinite(g1).all(axis=(1, 2)) == ok).all(), s
    num = np.linalg.norm((g1 - g2)[ok], axis=2)
    den = np.maximum(np.linalg.norm(g2[ok], axis=2), 1e-30)
    tol = 1e-5 + 1e-13 * cond[ok][:, None]
    say("snapshot", s, "max rel", (num / den).max(), "worst ratio to tolerance", (num / den / tol).max())
    assert (num / den <= tol).all(), s


def _single(pos, f, k):
    from meshoid_amd import engine

    ix = engine.SpatialIndex()
    ix.build(pos)
    out = engine.knn_hsml(ix, k, keep_dist=True)
    d1, _, st = engine.apply_D(out["nlist"], f)
    ngb, _ = out["nlist"].export()
    cond = moment_cond(pos.cpu().numpy(), ngb.cpu().numpy(), out["h"].cpu().numpy())
    return out["h"], out["rho"], d1, out["status"] | st, cond


def test_batch_matches_single_path():
    torch = _torch()
    from meshoid_amd.batch import SnapshotBatch

    sizes = [5000, 1200, 20000, 333, 9000, 64]
    ks = [32, 64, 32, 64, 32, 16]
    K = 5
    pos, fld = _snapshots(torch, sizes, K)
    say("batch eager", sizes)
    B = SnapshotBatch(sizes, ks, K, nbranch=3)
    o = B.run(pos, fld, graph=False)
    info = B.info()
    say("batch info", info)
    assert info["choice"] in info["candidates_ms"]
    for s in range(len(sizes)):
        h, rho, d1, st, cond = _single(pos[s], fld[s], ks[s])
        assert same(o["h"][s], h), s
        assert same(o["rho"][s], rho), s
        assert same(o["status"][s], st), s
        close(o["grad"][s], d1, s, cond)


def test_batch_wide_contraction():
    """K >= 8 takes the thread-group-per-query kernel: the same operator summed
    in another order, so float64 rounding differences only."""
    torch = _torch()
    from meshoid_amd.batch import SnapshotBatch

    sizes = [6000, 15000, 700]
    ks = [32, 64, 32]
    K = 12
    pos, fld = _snapshots(torch, sizes, K, seed=21)
    say("batch wide", sizes)
    B = SnapshotBatch(sizes, ks, K, nbranch=2)
    o = B.run(pos, fld, graph=False)
    info = B.info()
    say("batch info", info)
    assert info["kernel"] == "wide" and info["choice"] in (16, 32, 64)
    for s in range(len(sizes)):
        h, rho, d1, st, cond = _single(pos[s], fld[s], ks[s])
        assert same(o["h"][s], h) and same(o["rho"][s], rho) and same(o["status"][s], st), s
        close(o["grad"][s], d1, s, cond)


def test_batch_graph_replays():
    torch = _torch()
    from meshoid_amd.batch import SnapshotBatch

    sizes = [7000, 2500, 12000, 900]
    ks = [32, 64, 32, 64]
    K = 3
    pos, fld = _snapshots(torch, sizes, K, seed=5)
    B = SnapshotBatch(sizes, ks, K, nbranch=2)
    say("graph: eager reference")
    ref = B.run(pos, fld, graph=False)
    ref = {k: [t.clone() for t in v] for k, v in ref.items()}
    for rep in range(3):
        say("graph replay", rep)
        # eager work of the program between replays
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        for k in ref:
            for t in B.out[k]:
                t.fill_(-7) if t.dtype != torch.int32 else t.fill_(-1)
        o = B.run(pos, fld, graph=True)
        torch.cuda.synchronize()
        for k in ref:
            for s in range(len(sizes)):
                assert same(o[k][s], ref[k][s]), (rep, k, s)
        assert torch.isfinite(junk)
    # new input buffers: recaptured, same results
    say("graph recapture")
    pos2 = [p.clone() for p in pos]
    o = B.run(pos2, fld, graph=True)
    for s in range(len(sizes)):
        assert same(o["grad"][s], ref["grad"][s])


def test_batch_rejects_bad_input():
    torch = _torch()
    from meshoid_amd.batch import SnapshotBatch

    with pytest.raises(ValueError):
        SnapshotBatch([100, 10], [32, 32], 2)  # des_ngb above n
    with pytest.raises(NotImplementedError):
        SnapshotBatch([100, 200], [32, 200], 2)  # des_ngb above 128
    pos, fld = _snapshots(torch, [3000, 2000], 2, seed=9)
    B = SnapshotBatch([3000, 2000], 32, 2, nbranch=2)
    with pytest.raises(ValueError):
        B.run([pos[0], pos[1][:1000]], fld, graph=False)  # shape differs from the layout
