"""GPU forecaster network (forward, Poisson-NLL gradient, device Adam) against the FP64
numpy oracle (oracle/mlp.py, itself pinned by finite differences).

Bar: forward rate within 2e-5 relative; loss within 1e-6 relative; gradient within 1e-4
of the oracle's in norm (FP32 activations, FP64 sums over samples)."""
import numpy as np
import pytest

from gplum_amd import gpu
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def synthetic(n, seed):
    """Features in the ranges the system feeds (standardised logs) and Poisson counts
    from a smooth rate law."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 1.0, (n, 4))
    rate = np.exp(1.2 + 0.8 * x[:, 0] - 0.3 * x[:, 1] + 0.2 * x[:, 2] * x[:, 3])
    y = rng.poisson(rate).astype(np.float64)
    return x.astype(np.float32).astype(np.float64), y, rate


@pytest.mark.parametrize("n", [1, 300, 5000])
def test_forward_and_gradient_match_oracle(n):
    x, y, _ = synthetic(n, seed=n)
    theta = omlp.init_params(seed=3).astype(np.float32).astype(np.float64)
    lam_g = gpu.mlp_forward(theta, x)
    lam_o = omlp.forward(theta, x)
    assert np.max(np.abs(lam_g - lam_o) / lam_o) < 2e-5
    loss_g, grad_g = gpu.mlp_loss_grad(theta, x, y)
    loss_o, grad_o = omlp.loss_and_grad(theta, x, y)
    print("n %d loss %.9f vs %.9f" % (n, loss_g, loss_o))
    assert abs(loss_g - loss_o) <= 1e-6 * abs(loss_o) + 1e-9
    err = np.linalg.norm(grad_g - grad_o) / np.linalg.norm(grad_o)
    print("grad rel err %.3e" % err)
    assert err < 1e-4


@pytest.mark.parametrize("n", [1, 300, 100003])
def test_predict_matches_forward_to_bf16(n):
    """The sizing forecast (hidden layer on MFMA with bf16 operands, hardware tanh) against
    the FP64 oracle's rates, for the initial network and one trained on a rate law:
    bf16 keeps 8 bits, so rates agree to about 1 % (bars: median 1e-2, max 5e-2)."""
    x, y, _ = synthetic(n, seed=n + 7)
    xt, yt, _ = synthetic(20000, seed=9)
    theta0 = omlp.init_params(seed=3).astype(np.float32)
    trained, _, _ = gpu.mlp_train(theta0, xt, yt, 150, lr=3e-3)
    for theta in (theta0, trained):
        th = np.asarray(theta, np.float32).astype(np.float64)
        lam_p = gpu.mlp_predict(th, x)
        lam_o = omlp.forward(th, x)
        rel = np.abs(lam_p - lam_o) / lam_o
        print("n %d: predict vs oracle relative error median %.2e max %.2e" % (n, np.median(rel), rel.max()))
        assert np.median(rel) < 1e-2 and rel.max() < 5e-2


def test_gradient_is_deterministic():
    x, y, _ = synthetic(20000, seed=11)
    theta = omlp.init_params(seed=5)
    _, g1 = gpu.mlp_loss_grad(theta, x, y)
    _, g2 = gpu.mlp_loss_grad(theta, x, y)
    assert np.array_equal(g1, g2)


def test_device_adam_matches_host_adam():
    """Five device Adam steps equal the same Adam recursion run on the host with the
    oracle's gradients (to FP32 round-off), and training lowers the loss."""
    x, y, _ = synthetic(4000, seed=2)
    theta0 = omlp.init_params(seed=1).astype(np.float32)
    th_g, l0, l1 = gpu.mlp_train(theta0, x, y, 5, lr=1e-3)
    th = theta0.astype(np.float64)
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    for t in range(1, 6):
        _, g = omlp.loss_and_grad(th, x, y)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        th = th - 1e-3 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    # Adam normalises each step by |g|, so FP32 noise in the smallest gradients moves
    # single parameters by up to lr: compare in norm over the whole vector
    err = np.linalg.norm(th_g - th) / np.linalg.norm(th - theta0)
    print("device vs host Adam: rel err of the update %.3e" % err)
    assert err < 0.03
    assert l1 < l0


def test_training_learns_rate_law():
    """400 device Adam steps close three quarters of the gap between the best constant rate and
    the generating rate law."""
    x, y, rate = synthetic(20000, seed=4)
    theta0 = omlp.init_params(seed=2).astype(np.float32)
    th, l0, l1 = gpu.mlp_train(theta0, x, y, 400, lr=3e-3)
    const = y.mean()
    l_const = np.mean(const - y * np.log(const + omlp.EPS))
    l_true = np.mean(rate - y * np.log(rate + omlp.EPS))
    print("loss init %.4f trained %.4f constant-rate %.4f true-rate %.4f" % (l0, l1, l_const, l_true))
    assert l1 < l_const - 0.75 * (l_const - l_true)
