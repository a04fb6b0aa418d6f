"""The MLP oracle's analytic gradient, pinned by central finite differences
(float64, BF16 roundings switched off so the loss is smooth)."""
import numpy as np

from oracle import mlp_oracle as mo


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    X = rng.normal(size=(64, 16))
    t = rng.normal(size=64)
    p = rng.normal(size=16 * 64 + 64 + 64 * 64 + 64 + 64 + 1) * 0.2
    mo.ROUND = False
    try:
        g, _ = mo.grad(X, t, p.astype(np.float64))
        idx = rng.choice(len(p), 60, replace=False)
        for k in idx:
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            lp = np.mean((mo.forward(X, pp)[0] - t) ** 2)
            lm = np.mean((mo.forward(X, pm)[0] - t) ** 2)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k]) < 1e-5 * (1 + abs(fd)), (k, fd, g[k])
    finally:
        mo.ROUND = True


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 + 2 ** -7 - 2 ** -7  # tie to even: 1.0
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
