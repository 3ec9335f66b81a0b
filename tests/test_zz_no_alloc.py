"""mhip_integrate_noalloc (no device allocation; partial sums in a static
device array) against mhip_integrate through the C ABI: bit for bit, on two
calls in a row, the second with a smaller input, and its size limit."""
import ctypes
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


def _call(fn, ctx, vol, f, out):
    import torch

    f2 = f.reshape(f.shape[0], -1)
    return fn(ctx, ctypes.c_void_p(vol.data_ptr()), int(vol.dtype == torch.float64),
              ctypes.c_void_p(f2.data_ptr()), int(f2.dtype == torch.float64), f2.shape[0], f2.shape[1],
              ctypes.c_void_p(out.data_ptr()))


@pytest.mark.parametrize("vdt,fdt", [("float32", "float32"), ("float64", "float64"), ("float64", "float32")])
def test_integrate_noalloc_matches_integrate(vdt, fdt):
    torch = _torch()
    from meshoid_amd import engine
    from meshoid_amd._lib import lib

    ctx = engine.context(torch.device("cuda", 0))
    g = torch.Generator(device="cuda").manual_seed(5)
    for n, K in ((1_000_003, 7), (40_000, 3)):  # the second call is smaller
        say("integrate_noalloc n", n, "K", K, vdt, fdt)
        vol = torch.rand(n, device="cuda", generator=g).to(getattr(torch, vdt))
        f = torch.randn((n, K), device="cuda", generator=g).to(getattr(torch, fdt))
        want = torch.empty(K, dtype=torch.float64, device="cuda")
        got = torch.full((K,), float("nan"), dtype=torch.float64, device="cuda")
        assert _call(lib.mhip_integrate, ctx, vol, f, want) == 0
        assert _call(lib.mhip_integrate_noalloc, ctx, vol, f, got) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), (got, want)
        ref = (vol.double()[:, None] * f.double()).sum(0)
        torch.testing.assert_close(got, ref, rtol=1e-12, atol=1e-9)


def test_integrate_noalloc_rejects_too_many_components():
    torch = _torch()
    from meshoid_amd import engine
    from meshoid_amd._lib import NOALLOC_INTEGRATE_MAX_K, lib

    ctx = engine.context(torch.device("cuda", 0))
    K = NOALLOC_INTEGRATE_MAX_K + 1
    vol = torch.ones(100, dtype=torch.float64, device="cuda")
    f = torch.ones((100, K), dtype=torch.float64, device="cuda")
    out = torch.zeros(K, dtype=torch.float64, device="cuda")
    assert _call(lib.mhip_integrate_noalloc, ctx, vol, f, out) == -1
    assert b"at most" in lib.mhip_last_error(ctx)
    with pytest.raises(ValueError):
        engine.integrate_noalloc(vol, f, out)
    # the limit itself is accepted
    f = f[:, :NOALLOC_INTEGRATE_MAX_K].contiguous()
    out = torch.zeros(NOALLOC_INTEGRATE_MAX_K, dtype=torch.float64, device="cuda")
    engine.integrate_noalloc(vol, f, out)
    assert torch.equal(out, torch.full_like(out, 100.0))
