"""Pin the MLP oracle: its analytic gradient equals central finite differences of its loss,
and a few hundred full-batch gradient steps reduce the loss on a Poisson-count problem."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(37, 4))
    y = rng.poisson(1.5, 37).astype(float)
    theta = mlp.init_params(1)
    loss, g = mlp.loss_and_grad(theta, x, y)
    idx = rng.choice(mlp.NPARAM, 60, replace=False)
    for k in list(idx) + [mlp.NPARAM - 1, 0]:
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (mlp.loss_and_grad(tp, x, y)[0] - mlp.loss_and_grad(tm, x, y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_training_reduces_loss():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(512, 4))
    lam_true = np.exp(0.5 * x[:, 0] - 0.3 * x[:, 1])
    y = rng.poisson(lam_true).astype(float)
    theta = mlp.init_params(2)
    l0 = mlp.loss_and_grad(theta, x, y)[0]
    for _ in range(300):
        l, g = mlp.loss_and_grad(theta, x, y)
        theta -= 0.05 * g
    assert l < 0.8 * l0
