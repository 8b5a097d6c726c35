"""The warm-start network oracle (oracle/mlp.py) against finite differences
and exact BF16 rounding cases."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_cases():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    r = mlp.bf16_round(x)
    # 1 + 2^-8 is representable (8 bits of mantissa incl. hidden: step 2^-7)?
    # bf16 keeps 7 explicit mantissa bits: spacing at 1 is 2^-7
    assert r[0] == 1.0
    assert r[1] == 1.0                      # 2^-8 = half a step -> ties to even (1.0)
    assert r[2] == 1.0 + 2 ** -7           # 1.5 half-steps -> rounds up
    assert r[3] == 1.0                      # quarter step -> down
    assert r[4] == -2.5 and r[5] == 0.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nh = 16
    n = mlp.NIN * nh + nh + nh * nh + nh + nh * mlp.NOUT + mlp.NOUT
    p = rng.normal(size=n) * 0.3
    X = rng.normal(size=(9, mlp.NIN))
    T = rng.normal(size=(9, mlp.NOUT))
    loss, g = mlp.loss_and_grad(p, X, T, nh, quantise=False)
    for k in rng.choice(n, 40, replace=False):
        e = np.zeros(n); e[k] = 1e-6
        lp, _ = mlp.loss_and_grad(p + e, X, T, nh, quantise=False)
        lm, _ = mlp.loss_and_grad(p - e, X, T, nh, quantise=False)
        assert abs((lp - lm) / 2e-6 - g[k]) <= 1e-6 * max(1.0, abs(g[k]))
