"""The index's sort: a stable Hilbert order of the positions (the key
arithmetic restated on the host in float32, as the device computes it), on
the bits the index sorts (3D: all but the lowest 8 of the 63)."""
import sys

import numpy as np
import pytest

from hilbert_host import SORT_LO_BIT, frame_cells, hilbert_keys

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def host_keys(pos, bbox, dim=3):
    """Hilbert keys of the cells of the bounding box's cubic frame (restated
    on the host in float32 as the device forms them), without the low bits
    the index leaves unsorted."""
    return hilbert_keys(frame_cells(pos, bbox, dim), dim) >> np.uint64(SORT_LO_BIT[dim])


@pytest.mark.parametrize("n", [5000, 300000, 3000000])
def test_sort_is_stable_hilbert_order(n):
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
    # the bounding box exactly (16-byte vector reads from 2^18 particles on)
    lo, hi = np.asarray(ix.bbox()[:3], np.float32), np.asarray(ix.bbox()[3:6], np.float32)
    assert np.array_equal(lo, pos.min(0)) and np.array_equal(hi, pos.max(0))
    keys = host_keys(pos, ix.bbox())[perm]
    assert (keys[1:] >= keys[:-1]).all()
    tie = keys[1:] == keys[:-1]
    assert (perm[1:][tie] > perm[:-1][tie]).all()


@pytest.mark.parametrize("dim", [1, 2])
def test_sort_order_low_dimensions(dim):
    """1D and 2D keys pack dim*21 bits (every bit is sorted on)."""
    torch = _torch()
    from meshoid_amd import engine

    print("index order dim", dim, flush=True)
    rng = np.random.default_rng(dim)
    pos = rng.random((200000, dim)).astype(np.float32) * np.float32(7.0)
    ix = engine.SpatialIndex().build(torch.from_numpy(pos).cuda())
    perm = ix.order().cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(len(pos)))
    keys = host_keys(pos, ix.bbox(), dim)[perm]
    assert (keys[1:] >= keys[:-1]).all()


@pytest.mark.parametrize("where", ["middle", "tail"])
def test_nan_position_rejected_large(where):
    """A NaN coordinate fails the build also where the box is read in 16-byte
    vectors (n = 300001: the last particle's coordinates are the 3 floats after
    the last whole vector)."""
    torch = _torch()
    from meshoid_amd import engine

    print("nan", where, flush=True)
    n = 300001
    pos = np.random.default_rng(5).random((n, 3)).astype(np.float32)
    if where == "middle":
        pos[123457, 1] = np.nan
    else:
        pos[n - 1, 2] = np.nan
    with pytest.raises(ValueError):
        engine.SpatialIndex().build(torch.from_numpy(pos).cuda())
