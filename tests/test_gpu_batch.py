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
    assert info["choice"] in info["candidates"] and info["choice"]["kernel"] == "tile"
    assert 1 <= info["branches"] <= 3
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
    assert info["choice"]["kernel"] == "wide" and info["choice"]["param"] in (8, 16, 32)
    assert [c["kernel"] for c in info["candidates"]] == ["wide"] * 3  # K = 12: no matrix-core kernel
    for s in range(len(sizes)):
        h, rho, d1, st, cond = _single(pos[s], fld[s], ks[s])
        assert same(o["h"][s], h) and same(o["rho"][s], rho) and same(o["status"][s], st), s
        close(o["grad"][s], d1, s, cond)


@pytest.mark.parametrize("K,kmax", [(12, 64), (32, 64), (16, 128)])
@pytest.mark.parametrize("cand", [0, 1, 2])
def test_batch_wide_candidates(K, kmax, cand):
    """Each thread-group width (8, 16, 32 threads per query, summing 4, 2 or
    1 fields per pass; K = 12 leaves part of the last pass idle), forced,
    against the single-set float64 path; des_ngb up to 128 sizes the LDS
    for the largest lists (57 KB per block of 16 queries)."""
    torch = _torch()
    from meshoid_amd.batch import SnapshotBatch

    sizes = [6000, 15000, 700]
    ks = [32, kmax, 32]
    pos, fld = _snapshots(torch, sizes, K, seed=51 + K)
    say("batch wide candidate", cand, "K", K)
    B = SnapshotBatch(sizes, ks, K, nbranch=2)
    B.force(candidate=cand, nbranch=2)
    o = B.run(pos, fld, graph=False)
    info = B.info()
    assert info["choice"]["kernel"] == "wide" and info["choice"]["param"] == (8, 16, 32)[cand]
    for s in range(len(sizes)):
        h, rho, d1, st, cond = _single(pos[s], fld[s], ks[s])
        assert same(o["status"][s], st), s
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


@pytest.mark.parametrize("K", [16, 32])
def test_batch_mfma_contraction(K):
    """The matrix-core contraction (forced, so it is checked whatever the
    timing picks) against the single-set float64 path.  It multiplies and
    sums in float32, so each output component may differ by the float32
    rounding of its k terms: |err| <= (k + 4) 2^-24 sum_j |g_j| |f_j - f_i|,
    with g_j the oracle's weights on the same neighbours (a wrong lane layout
    gives errors of the order of the result itself)."""
    torch = _torch()
    from meshoid_amd import engine
    from meshoid_amd.batch import SnapshotBatch
    from oracle import oracle as orc

    sizes = [6000, 15000, 700]
    ks = [32, 64, 32]
    pos, fld = _snapshots(torch, sizes, K, seed=31 + K)
    say("batch mfma", K, sizes)
    B = SnapshotBatch(sizes, ks, K, nbranch=2)
    B.force(candidate=3, nbranch=2)
    o = B.run(pos, fld, graph=False)
    info = B.info()
    say("batch info", info)
    assert info["choice"]["kernel"] == "mfma" and info["choice"]["param"] == K and info["branches"] == 2
    rng = np.random.default_rng(K)
    for s in range(len(sizes)):
        h, rho, d1, st, cond = _single(pos[s], fld[s], ks[s])
        assert same(o["h"][s], h) and same(o["rho"][s], rho) and same(o["status"][s], st), s
        ix = engine.SpatialIndex().build(pos[s])
        out = engine.knn_hsml(ix, ks[s], keep_dist=True)
        ngb = out["nlist"].export()[0].cpu().numpy().astype(np.int64)
        h64 = out["nlist"].h64().cpu().numpy()
        p = pos[s].cpu().numpy().astype(np.float64)
        f = fld[s].cpu().numpy().astype(np.float64)
        idx = np.sort(rng.choice(sizes[s], min(1500, sizes[s]), replace=False))
        G = orc.gradient_weights(p, ngb[idx], h64[idx], idx)  # (m, 3, k)
        df = np.abs(f[ngb[idx]] - f[idx][:, None, :])  # (m, k, K)
        S = np.einsum("mcj,mjf->mfc", np.abs(G), df)  # (m, K, 3)
        g1 = o["grad"][s].cpu().numpy().astype(np.float64)[idx]
        g2 = d1.cpu().numpy().astype(np.float64)[idx]
        ok = np.isfinite(g2).all(axis=(1, 2))
        assert (np.isfinite(g1).all(axis=(1, 2)) == ok).all(), s
        bound = (ks[s] + 4) * 2.0 ** -24 * S + 1e-7 * np.abs(g2)
        err = np.abs(g1 - g2)
        say("snapshot", s, "max err / bound", (err[ok] / bound[ok]).max())
        assert (err[ok] <= bound[ok]).all(), s


def test_batch_branch_counts_give_identical_results():
    """The branch streams only reorder independent work: 1, 2 or 4 branches
    give bit-identical results."""
    torch = _torch()
    from meshoid_amd.batch import SnapshotBatch

    sizes = [4000, 9000, 1500, 6000, 300]
    ks = [32, 64, 32, 32, 16]
    K = 8
    pos, fld = _snapshots(torch, sizes, K, seed=41)
    res = []
    for nb in (1, 2, 4):
        say("branches", nb)
        B = SnapshotBatch(sizes, ks, K, nbranch=4)
        B.force(candidate=0, nbranch=nb)
        o = B.run(pos, fld, graph=False)
        assert B.info()["branches"] == nb
        res.append({k: [t.clone() for t in v] for k, v in o.items()})
    for r in res[1:]:
        for k in r:
            for s in range(len(sizes)):
                assert same(r[k][s], res[0][k][s]), (k, s)
