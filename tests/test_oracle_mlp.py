"""The CPU MLP restatement (oracle/mlp.py) against what the mathematics fixes:
central finite differences of its loss for the gradients, the closed forms of
GELU at known points, and Adam's first step (whose size is lr times the sign
of the gradient, since m / sqrt(v) = g / |g| after bias correction)."""
import numpy as np

from oracle import mlp as om


def _net(rng, dims):
    return [(rng.standard_normal((dims[i + 1], dims[i])) / np.sqrt(dims[i]), rng.standard_normal(dims[i + 1]) * 0.1)
            for i in range(len(dims) - 1)]


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 3.0])
    # Phi(1) = 0.8413447460685429, Phi(3) = 0.9986501019683699
    assert np.allclose(om.gelu(z), [0.0, 0.8413447460685429, -0.15865525393145707, 3 * 0.9986501019683699], atol=1e-15)
    h = 1e-6
    assert np.allclose(om.gelu_d(z), (om.gelu(z + h) - om.gelu(z - h)) / (2 * h), atol=1e-8)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    dims = [5, 7, 6, 3]
    p = _net(rng, dims)
    x = rng.standard_normal((9, 5))
    t = rng.standard_normal((9, 3))
    g, _ = om.grads(p, x, t)
    h = 1e-6
    for li, (W, b) in enumerate(p):
        for arr, garr in ((W, g[li][0]), (b, g[li][1])):
            it = np.nditer(arr, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = arr[idx]
                arr[idx] = old + h
                lp = om.loss(p, x, t)
                arr[idx] = old - h
                lm = om.loss(p, x, t)
                arr[idx] = old
                assert abs((lp - lm) / (2 * h) - garr[idx]) <= 1e-7 * max(1.0, abs(garr[idx])), (li, idx)


def test_adam_first_step_is_lr_times_sign():
    rng = np.random.default_rng(4)
    p = _net(rng, [4, 3])
    g, _ = om.grads(p, rng.standard_normal((6, 4)), rng.standard_normal((6, 3)))
    st = [[np.zeros_like(p[0][0]), np.zeros_like(p[0][0]), np.zeros_like(p[0][1]), np.zeros_like(p[0][1])]]
    new = om.adam_step(p, g, st, 1, 1e-3)
    step = p[0][0] - new[0][0]
    assert np.allclose(step, 1e-3 * np.sign(g[0][0]) * np.abs(g[0][0]) / (np.abs(g[0][0]) + 1e-8), rtol=1e-12)
