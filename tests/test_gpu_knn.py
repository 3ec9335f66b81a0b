"""GPU parity: exact kNN, smoothing length and density through the C ABI,
against the FP64 oracle and the committed cKDTree golden vectors."""
import glob
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def set_mismatch(ngb_g, d_g, ngb_o, d_o, rtol=1e-6):
    """Number of rows whose sets differ beyond float64 ties at the k-th
    distance (allowing the float32 rounding of the reported distance)."""
    bad = 0
    for i in range(ngb_o.shape[0]):
        a, b = set(ngb_g[i].tolist()), set(ngb_o[i].tolist())
        if a == b:
            continue
        kth = d_o[i, -1]
        dd = dict(zip(ngb_o[i].tolist(), d_o[i].tolist()))
        for j in a ^ b:
            if j in dd and abs(dd[j] - kth) <= 4.5e-16 * kth:
                continue
            if j not in dd and abs(d_g[i][list(ngb_g[i]).index(j)] - kth) <= rtol * kth:
                # member only on the GPU side: must tie the oracle's k-th in fp64
                continue
            bad += 1
            break
    return bad


def gpu_knn(pos, k, box=None, qidx=None):
    torch = _torch()
    from meshoid_amd import engine

    p = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    ix = engine.SpatialIndex().build(p, box)
    q = None if qidx is None else torch.from_numpy(np.asarray(qidx, dtype=np.int64)).cuda()
    out = engine.knn_hsml(ix, k, qidx=q, keep_dist=True)
    ngb, dist = out["nlist"].export()
    torch.cuda.synchronize()
    return (ngb.cpu().numpy().astype(np.int64), dist.cpu().numpy().astype(np.float64),
            out["h"].cpu().numpy().astype(np.float64), out["status"].cpu().numpy(), out)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "knn_*.npz"))))
def test_knn_matches_ckdtree_golden(path):
    z = np.load(path)
    box = float(z["boxsize"])
    k = int(z["k"])
    say("case", os.path.basename(path), "k", k, "box", box)
    ngb, dist, h, st, _ = gpu_knn(z["pos"], k, box if box > 0 else None, qidx=z["qidx"])
    assert set_mismatch(ngb, dist, z["ngb"].astype(np.int64), z["dist"]) == 0
    np.testing.assert_allclose(dist, z["dist"], rtol=2e-6, atol=1e-12)
    assert (ngb[:, 0] == z["qidx"]).all()
    assert (st & 16 == 0).all()


CASES = [
    # name, n, dim, k, box, dtype, kind
    ("uniform_f64_k32", 20000, 3, 32, None, np.float64, "uniform"),
    ("uniform_f32_k32", 20000, 3, 32, None, np.float32, "uniform"),
    ("periodic_f32_k32", 20000, 3, 32, 1.0, np.float32, "uniform"),
    ("periodic_f64_k64", 20000, 3, 64, 1.0, np.float64, "uniform"),
    ("clustered_f32_k32", 40000, 3, 32, None, np.float32, "clustered"),
    ("clustered_f32_k8", 40000, 3, 8, None, np.float32, "clustered"),
    ("uniform2d_f64_k20", 20000, 2, 20, None, np.float64, "uniform"),
    ("uniform1d_f64_k4", 5000, 1, 4, None, np.float64, "uniform"),
    ("uniform_f64_k100", 8000, 3, 100, 1.0, np.float64, "uniform"),
    ("uniform_f32_k128", 8000, 3, 128, None, np.float32, "uniform"),
    ("tiny_n_eq_k", 33, 3, 32, None, np.float64, "uniform"),
]


def make(n, dim, dtype, kind, seed=0):
    from meshoid_amd import synthetic

    if kind == "uniform":
        return synthetic.uniform(n, dim, seed=seed, dtype=dtype)
    return synthetic.clustered(n, seed=seed, nclump=32, background=0.1, dtype=dtype)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_knn_and_hsml_match_oracle(oracle, case):
    name, n, dim, k, box, dtype, kind = case
    say("case", name)
    pos = make(n, dim, dtype, kind)
    ngb, dist, h, st, out = gpu_knn(pos, k, box)
    d_o, n_o = oracle.knn(pos.astype(np.float64), k, box)
    assert set_mismatch(ngb, dist, n_o, d_o) == 0
    np.testing.assert_allclose(dist, d_o, rtol=2e-6, atol=1e-12)
    h_o = oracle.hsml_iter(d_o, dim=dim, error_norm=1e-13)
    rel = np.abs(h / h_o - 1)
    say("max rel h err", rel.max(), "ambiguous rows", int((st & 1).sum()))
    assert rel.max() <= 1e-6
    rho = out["rho"].cpu().numpy().astype(np.float64)
    vn = {1: 2.0, 2: np.pi, 3: 4 * np.pi / 3}[dim]
    np.testing.assert_allclose(rho, k / (vn * h_o**dim), rtol=4e-6)
    np.testing.assert_allclose(out["vol"].cpu().numpy(), vn * h_o**dim / k, rtol=4e-6)
    assert (st & (16 | 4) == 0).all()


def test_knn_masked_rows_and_masses(oracle):
    torch = _torch()
    from meshoid_amd import engine

    pos = make(30000, 3, np.float32, "clustered", seed=4)
    rng = np.random.default_rng(0)
    qidx = rng.choice(30000, 5000, replace=True)  # arbitrary order, repeats allowed
    m = rng.random(5000) + 0.5
    p = torch.from_numpy(pos).cuda()
    ix = engine.SpatialIndex().build(p)
    out = engine.knn_hsml(ix, 32, qidx=torch.from_numpy(qidx).cuda(),
                          m_rows=torch.from_numpy(m).cuda())
    ngb, dist = out["nlist"].export()
    d_o, n_o = oracle.knn(pos.astype(np.float64), 32, queries=pos[qidx].astype(np.float64))
    assert set_mismatch(ngb.cpu().numpy(), dist.cpu().numpy(), n_o, d_o) == 0
    h_o = oracle.hsml_iter(d_o, error_norm=1e-13)
    h = out["h"].cpu().numpy().astype(np.float64)
    assert np.abs(h / h_o - 1).max() <= 1e-6
    np.testing.assert_allclose(out["rho"].cpu().numpy(), 32 * m / (4 * np.pi / 3 * h_o**3),
                               rtol=4e-6)


def test_knn_duplicates_and_lattice_ties(oracle):
    """Exact float64 ties: sets may differ only among members tied with the k-th."""
    g = np.arange(12, dtype=np.float64)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = np.concatenate([lat, lat[:100]])  # 100 duplicated points
    ngb, dist, h, st, _ = gpu_knn(pos, 32)
    d_o, n_o = oracle.knn(pos, 32)
    assert set_mismatch(ngb, dist, n_o, d_o) == 0
    np.testing.assert_allclose(dist, d_o, rtol=1e-6, atol=1e-12)
    h_o = oracle.hsml_iter(d_o, error_norm=1e-13)
    assert np.abs(h / h_o - 1).max() <= 1e-6


def test_knn_band_overflow_ties(oracle):
    """More float64 ties at the k-th distance than the resolver's band holds:
    8 copies of every point of a 6^3 lattice give an interior particle 8
    members at distance 0 and 48 candidates tied at distance 1 for the
    remaining 24 places.  The overflow is flagged, and the sets and distances
    still agree with the oracle up to those ties."""
    from meshoid_amd._lib import ST_BAND_OVERFLOW

    g = np.arange(6, dtype=np.float64)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = np.repeat(lat, 8, axis=0)
    say("band overflow", len(pos))
    ngb, dist, h, st, _ = gpu_knn(pos, 32)
    d_o, n_o = oracle.knn(pos, 32)
    interior = np.all((lat > 0) & (lat < 5), axis=1).repeat(8)
    assert (st[interior] & ST_BAND_OVERFLOW).all()
    assert set_mismatch(ngb, dist, n_o, d_o) == 0
    np.testing.assert_allclose(dist, d_o, rtol=1e-6, atol=1e-12)
    # every row holds its own copies and 24 particles at distance 1.  (h is
    # not compared: 8 coincident particles alone weigh (32/3) 8 > 32, so the
    # smoothing-length equation has no root and both solvers stop at the
    # bottom of their brackets.)
    assert (np.sum(dist[interior] == 0.0, axis=1) == 8).all()
    assert (np.sum(dist[interior] == 1.0, axis=1) == 24).all()


def test_knn_errors():
    torch = _torch()
    from meshoid_amd import engine

    p = torch.rand(100, 3, device="cuda")
    with pytest.raises(ValueError):
        engine.SpatialIndex().build(p * 2.0, boxsize=1.0)  # outside the periodic box
    bad = p.clone()
    bad[5, 1] = float("nan")
    with pytest.raises(ValueError):
        engine.SpatialIndex().build(bad)
    ix = engine.SpatialIndex().build(p)
    with pytest.raises(ValueError):
        engine.knn_hsml(ix, 101)
    with pytest.raises(NotImplementedError):
        engine.knn_hsml(engine.SpatialIndex().build(torch.rand(300, 3, device="cuda")), 200)


def test_knn_large_clustered_invariants():
    """2^22 clustered particles: size-independent properties at scale."""
    torch = _torch()
    from meshoid_amd import engine, synthetic

    pos = synthetic.clustered(1 << 22, seed=9, nclump=256, background=0.1)
    p = torch.from_numpy(pos).cuda()
    ix = engine.SpatialIndex().build(p)
    out = engine.knn_hsml(ix, 32, keep_dist=True)
    ngb, dist = out["nlist"].export()
    st = out["status"]
    assert int(((st & 16) != 0).sum()) == 0
    assert bool((ngb[:, 0].long() == torch.arange(1 << 22, device="cuda")).all() |
                (dist[:, 0] == 0).all())
    assert bool((dist[:, 1:] >= dist[:, :-1]).all())
    # the kernel sum at h equals des_ngb: over the float32 distances and h to
    # float32 rounding, and over the float64 distances with the list's
    # float64-refined h to ~1e-12
    def nsum(d, h):
        q = d / h[:, None]
        W = torch.where(q <= 0.5, 1 - 6 * q**2 + 6 * q**3, torch.where(q <= 1, 2 * (1 - q) ** 3, 0 * q))
        return 32 / 3 * W.sum(1)

    assert float((nsum(dist.double(), out["h"].double()) - 32).abs().max()) < 1e-3
    _, d64 = out["nlist"].export(want_ngb=False, f64=True)
    h64 = out["nlist"].h64()
    err = float((nsum(d64, h64) - 32).abs().max())
    say("max |n(h) - 32| with float64 distances and h:", err)
    assert err < 1e-9
    say("ambiguous rows resolved in fp64:", int(((st & 1) != 0).sum()))


@pytest.mark.parametrize("case", ["clustered_f32", "periodic_f64", "masked", "k64"])
def test_leaf_record_overflow_path_matches(case):
    """Wavefronts whose leaf record overflows collect their members by a
    second tree walk with the final radii.  The test hook caps the record at
    8 leaves, so every wavefront takes that path; neighbour sets, distances,
    smoothing lengths and densities must equal the normal path's."""
    torch = _torch()
    from meshoid_amd import engine, synthetic

    say("overflow path", case)
    k = 64 if case == "k64" else 32
    box = None
    if case == "periodic_f64":
        pos = np.random.default_rng(3).random((60000, 3)) * 10.0
        box = 10.0
    else:
        pos = synthetic.clustered(200000, seed=11, nclump=32, box=10.0, background=0.1)
    qidx = None
    if case == "masked":
        qidx = np.sort(np.random.default_rng(5).choice(len(pos), 30000, replace=False))
    p = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    ix = engine.SpatialIndex().build(p, box)
    q = None if qidx is None else torch.from_numpy(qidx.astype(np.int64)).cuda()
    res = []
    for small in (False, True):
        out = engine.knn_hsml(ix, k, qidx=q, keep_dist=True, small_record=small)
        ngb, dist = out["nlist"].export()
        h64 = out["nlist"].h64()
        torch.cuda.synchronize()
        res.append((ngb.cpu().numpy(), dist.cpu().numpy(), h64.cpu().numpy(), out["rho"].cpu().numpy(),
                    out["status"].cpu().numpy()))
    (n0, d0, h0, r0, s0), (n1, d1, h1, r1, s1) = res
    assert np.array_equal(n0, n1)
    assert np.array_equal(d0, d1)
    assert np.array_equal(s0, s1)
    # the float64 refinement sums the same members in another order
    np.testing.assert_allclose(h1, h0, rtol=1e-14, atol=0)
    np.testing.assert_allclose(r1, r0, rtol=3e-7, atol=0)
