"""Oracle of the neighbour-count forecaster: a 4-64-64-1 MLP (tanh hidden layers, softplus
Poisson rate) with the Poisson negative log-likelihood, written in numpy FP64.

Test infrastructure only: it checks the HIP forward/backward kernels (gplum_amd/csrc/mlp.hip)
and is itself pinned by central finite differences in tests/test_mlp_oracle.py.

Parameter vector layout (shared with the HIP kernels, float32 there):
    W1[64][4], b1[64], W2[64][64], b2[64], w3[64], b3[1]   (row-major, 4545 values)
"""
import numpy as np

NIN, NH = 4, 64
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, 1 / np.sqrt(NIN), (NH, NIN)), "b1": np.zeros(NH),
         "W2": rng.normal(0, 1 / np.sqrt(NH), (NH, NH)), "b2": np.zeros(NH),
         "w3": rng.normal(0, 1 / np.sqrt(NH), NH), "b3": np.array([-2.0])}
    return np.concatenate([p[n].ravel() for n, _ in SIZES])


def softplus(z):
    return np.where(z > 20, z, np.log1p(np.exp(np.minimum(z, 20))))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


EPS = 1e-6


def forward(theta, x):
    """Poisson rate lambda for features x[n,4]."""
    P = unpack(np.asarray(theta, np.float64))
    h1 = np.tanh(x @ P["W1"].T + P["b1"])
    h2 = np.tanh(h1 @ P["W2"].T + P["b2"])
    z = h2 @ P["w3"] + P["b3"][0]
    return softplus(z)


def loss_and_grad(theta, x, y):
    """Mean Poisson NLL  mean(lambda - y log(lambda + EPS))  and its gradient."""
    P = unpack(np.asarray(theta, np.float64))
    n = len(x)
    a1 = x @ P["W1"].T + P["b1"]
    h1 = np.tanh(a1)
    a2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.tanh(a2)
    z = h2 @ P["w3"] + P["b3"][0]
    lam = softplus(z)
    loss = np.mean(lam - y * np.log(lam + EPS))
    dz = (1.0 - y / (lam + EPS)) * sigmoid(z) / n
    d2 = np.outer(dz, P["w3"]) * (1.0 - h2 * h2)
    d1 = (d2 @ P["W2"]) * (1.0 - h1 * h1)
    g = {"W1": d1.T @ x, "b1": d1.sum(0), "W2": d2.T @ h1, "b2": d2.sum(0),
         "w3": h2.T @ dz, "b3": np.array([dz.sum()])}
    return loss, np.concatenate([g[nm].ravel() for nm, _ in SIZES])
