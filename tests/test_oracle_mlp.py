"""Pins the numpy restatement of the cost-model network (oracle/mlp.py): its analytic
gradient must equal central finite differences of its own loss in float64.  CPU only."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    F, H, M = 12, 16, 9
    rng = np.random.default_rng(0)
    n = F * H + H + H * H + H + H + 1
    params = rng.normal(scale=0.4, size=n)
    X = rng.uniform(size=(M, F))
    y = rng.normal(size=M)
    loss, g = mlp.loss_and_grad(params, X, y, F, H)
    eps = 1e-6
    idx = rng.choice(n, size=60, replace=False)
    for i in idx:
        p = params.copy()
        p[i] += eps
        lp, _ = mlp.loss_and_grad(p, X, y, F, H)
        p[i] -= 2 * eps
        lm, _ = mlp.loss_and_grad(p, X, y, F, H)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_planted_sign_error_is_caught():
    # a flipped sign in one gradient block must fail the finite-difference check
    F, H, M = 4, 8, 5
    rng = np.random.default_rng(1)
    n = F * H + H + H * H + H + H + 1
    params = rng.normal(scale=0.5, size=n)
    X = rng.uniform(size=(M, F))
    y = rng.normal(size=M)
    _, g = mlp.loss_and_grad(params, X, y, F, H)
    bad = g.copy()
    bad[F * H: F * H + H] *= -1  # b1 block
    eps = 1e-6
    worst = 0.0
    for i in range(F * H, F * H + H):
        p = params.copy()
        p[i] += eps
        lp, _ = mlp.loss_and_grad(p, X, y, F, H)
        p[i] -= 2 * eps
        lm, _ = mlp.loss_and_grad(p, X, y, F, H)
        worst = max(worst, abs((lp - lm) / (2 * eps) - bad[i]))
    assert worst > 1e-4
