"""radtransfer (src/meshoid/radiation/radtransfer.py:11-177) on the GPU
against the oracle's sequential farthest-first sweep
(tests/test_radtransfer_oracle.py pins it): slabs composed exactly, one or
several bands, scalar / per-band / per-pixel backgrounds, optically thick
and thin particles, particles off the image."""
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


@pytest.mark.parametrize("n,nf,res,i0", [(3000, 1, 32, 0.0), (8000, 3, 48, "band"), (20000, 2, 100, "pixel")])
def test_radtransfer_matches_oracle(oracle, n, nf, res, i0):
    _torch()
    import meshoid_amd

    rng = np.random.default_rng(n)
    x = rng.normal(0.0, 3.0, (n, 3))
    m = rng.random(n) + 0.1
    h = np.exp(rng.normal(-0.3, 0.5, n))
    j = rng.random((n, nf)) * 10
    kap = np.exp(rng.normal(0.0, 1.5, (n, nf)))  # optically thin to thick
    if i0 == "band":
        i0v = rng.random(nf)
    elif i0 == "pixel":
        i0v = rng.random((res, res, nf))
    else:
        i0v = i0
    say("radtransfer n", n, "bands", nf, "res", res, "i0", i0)
    img = meshoid_amd.radtransfer(j, m, kap, x, h, res, 16.0, center=[0.5, -0.3, 0.0], i0=i0v)
    want = oracle.radtransfer(j, m, kap, x, h, res, 16.0, center=[0.5, -0.3, 0.0], i0=i0v)
    assert img.shape == (res, res, nf)
    err = np.abs(img - want).max() / np.abs(want).max()
    say("  max rel to image max", err)
    assert err <= 1e-11
