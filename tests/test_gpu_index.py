"""The index's sort: a stable Morton order of the positions (the key
arithmetic restated on the host in float32, as the device computes it)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def spread21(v):
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    for s, m in ((32, 0x1F00000000FFFF), (16, 0x1F0000FF0000FF), (8, 0x100F00F00F00F00F),
                 (4, 0x10C30C30C30C30C3), (2, 0x1249249249249249)):
        v = (v | (v << np.uint64(s))) & np.uint64(m)
    return v


def host_keys(pos, bbox):
    """21-bit-per-axis Morton keys in the cubic frame of the bounding box:
    t = (x - lo) * (2097151 / longest side), float32, clamped, truncated."""
    lo = np.asarray(bbox[:3], np.float32)
    ext = np.float32(max(np.float32(bbox[3 + c]) - lo[c] for c in range(3)))
    scale = np.float32(2097151.0) / ext
    key = np.zeros(len(pos), np.uint64)
    for c in range(3):
        t = (pos[:, c].astype(np.float32) - lo[c]) * scale
        t = np.minimum(np.maximum(t, np.float32(0)), np.float32(2097151.0))
        key |= spread21(t.astype(np.uint64)) << np.uint64(c)
    return key


@pytest.mark.parametrize("n", [5000, 300000, 3000000])
def test_sort_is_stable_morton_order(n):
    torch = _torch()
    from meshoid_amd import engine, synthetic

    print("index order", n, flush=True)
    sys.stdout.flush()
    pos = synthetic.clustered(n, seed=n, nclump=64, box=10.0, background=0.1)
    # repeated positions: the order among equal keys must keep the input order
    pos[n // 2: n // 2 + 100] = pos[7]
    ix = engine.SpatialIndex().build(torch.from_numpy(pos).cuda())
    perm = ix.order().cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(n))
    keys = host_keys(pos, ix.bbox())[perm]
    assert (keys[1:] >= keys[:-1]).all()
    tie = keys[1:] == keys[:-1]
    assert (perm[1:][tie] > perm[:-1][tie]).all()
