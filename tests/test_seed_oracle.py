"""CPU pin of the seed-network oracle: its backward pass against central
finite differences of its own forward pass (float64, rounding disabled)."""
import numpy as np

from oracle import mlp_oracle as mo


def test_oracle_gradient_matches_finite_differences(monkeypatch):
    monkeypatch.setattr(mo, "bf16", lambda x: np.asarray(x, dtype=np.float64))
    rng = np.random.default_rng(0)
    P = rng.standard_normal(mo.NPARAM) * 0.2
    F = rng.standard_normal((40, mo.NIN))
    T = rng.standard_normal(40)
    mu, sd = np.zeros(mo.NIN), np.ones(mo.NIN)
    L, g = mo.loss_and_grad(P, F, T, mu, sd)
    idx = rng.choice(mo.NPARAM, 60, replace=False)
    for i in idx:
        e = np.zeros(mo.NPARAM)
        e[i] = 1e-6
        Lp, _ = mo.loss_and_grad(P + e, F, T, mu, sd)
        Lm, _ = mo.loss_and_grad(P - e, F, T, mu, sd)
        fd = (Lp - Lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert mo.NPARAM == 6081


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-9, 1.0 + 3 * 2**-9, -2.5, 0.0], dtype=np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0
    # 1 + 2^-8 is the midpoint of 1 and 1 + 2^-7: ties to the even mantissa (1)
    assert r[1] == 1.0
    # below / above the midpoint
    assert r[2] == 1.0
    assert r[3] == 1.0 + 2**-7
    assert r[4] == -2.5 and r[5] == 0.0
