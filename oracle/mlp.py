"""CPU FP64 restatement of the surrogate's MLP: forward pass, mean-squared-error
gradients and one Adam step.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the BF16 matrix-core
kernels of mpifdtd_amd (forward outputs, gradients, the Adam update), never by
the product package.  The network has no counterpart in the reference; this
file fixes the mathematics the kernels must reproduce:

    a_0 = x,  z_l = a_l W_l^T + b_l,  a_{l+1} = GELU(z_l) (hidden layers),
    y = z_{L-1} (linear output),  loss = mean over the B x d outputs of (y - t)^2,
    GELU(z) = z Phi(z),  GELU'(z) = Phi(z) + z phi(z),
    Adam (Kingma and Ba 2015): m = 0.9 m + 0.1 g, v = 0.999 v + 0.001 g^2,
    w -= lr (m / (1 - 0.9^t)) / (sqrt(v / (1 - 0.999^t)) + 1e-8).

Its gradients are pinned by central finite differences in
tests/test_oracle_mlp.py.
"""
import numpy as np
from scipy.special import erf


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, x):
    """params: list of (W [out, in], b [out]); returns y and the per-layer (a, z)."""
    a = np.asarray(x, np.float64)
    cache = []
    for i, (W, b) in enumerate(params):
        z = a @ np.asarray(W, np.float64).T + np.asarray(b, np.float64)
        cache.append((a, z))
        a = gelu(z) if i + 1 < len(params) else z
    return a, cache


def loss(params, x, t):
    y, _ = forward(params, x)
    return float(np.mean((y - t) ** 2))


def grads(params, x, t):
    """[(dW, db)] of the mean squared error, and the loss."""
    y, cache = forward(params, x)
    B, d = y.shape
    g = 2.0 * (y - t) / (B * d)
    out = [None] * len(params)
    for i in range(len(params) - 1, -1, -1):
        a, z = cache[i]
        out[i] = (g.T @ a, g.sum(axis=0))
        if i > 0:
            g = (g @ np.asarray(params[i][0], np.float64)) * gelu_d(cache[i - 1][1])
    return out, float(np.mean((y - t) ** 2))


def adam_step(params, gr, state, step, lr):
    """One Adam step in place on float64 copies; state: list of (mW, vW, mb, vb)."""
    c1, c2 = 1.0 - 0.9 ** step, 1.0 - 0.999 ** step
    new = []
    for (W, b), (dW, db), st in zip(params, gr, state):
        res = []
        for k, (w, g) in enumerate(((W, dW), (b, db))):
            m = 0.9 * st[2 * k] + 0.1 * g
            v = 0.999 * st[2 * k + 1] + 0.001 * g * g
            st[2 * k], st[2 * k + 1] = m, v
            res.append(np.asarray(w, np.float64) - lr * (m / c1) / (np.sqrt(v / c2) + 1e-8))
        new.append(tuple(res))
    return new
