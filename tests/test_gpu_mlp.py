"""The surrogate's BF16 matrix-core kernels against the CPU FP64 restatement
(oracle/mlp.py): the product in every operand layout, the forward pass, the
gradients of every layer, the Adam update, and a regression that the
training must fit.

Tolerances: a product of BF16 operands accumulated in FP32 is exact up to the
FP32 accumulation (1e-5 relative here, K <= 512); the network's activations
and gradients pass through BF16 between layers (8 significant bits), so the
forward pass and gradients are held to 3e-2 relative L2 against FP64.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _bf16_bits(a):
    """round-to-nearest-even FP32 -> BF16 bit patterns, and their values"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    b = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return b, (b.astype(np.uint32) << 16).view(np.float32)


def _mods():
    import mpifdtd_amd._native as N
    from mpifdtd_amd import surrogate as S
    from oracle import mlp as om
    return N, S, om


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
def test_gemm_layouts(ta, tb):
    N, S, om = _mods()
    rng = np.random.default_rng(10 + 2 * ta + tb)
    M, Nn, K = 128, 192, 96
    a_bits, a_val = _bf16_bits(rng.standard_normal((M, K)))
    b_bits, b_val = _bf16_bits(rng.standard_normal((K, Nn)))
    a_store = a_bits.T.copy() if ta else a_bits      # a[K][M] or a[M][K]
    b_store = b_bits.copy() if tb else b_bits.T.copy()  # b[K][N] or b[N][K]
    da = S.DeviceArray(a_store, dtype=np.uint16)
    db = S.DeviceArray(b_store, dtype=np.uint16)
    dc = S.DeviceArray(shape=(M, Nn))
    assert N.lib().mlp_gemm_bf16(da.ptr, db.ptr, dc.ptr, M, Nn, K, ta, tb) == 0
    ref = a_val.astype(np.float64) @ b_val.astype(np.float64)
    assert rel_l2(dc.numpy(), ref) <= 1e-5


def _net_and_data(S, dims, B, seed):
    rng = np.random.default_rng(seed)
    net = S.MLP(dims, batch_max=256, seed=seed)
    params = [((rng.standard_normal((dims[i + 1], dims[i])) / np.sqrt(dims[i])).astype(np.float32),
               (0.1 * rng.standard_normal(dims[i + 1])).astype(np.float32)) for i in range(len(dims) - 1)]
    for l, (w, b) in enumerate(params):
        net.set_params(l, w, b)
    x = rng.random((B, dims[0])).astype(np.float32)
    t = rng.standard_normal((B, dims[-1])).astype(np.float32)
    return net, params, x, t


def test_forward_and_gradients_against_fp64_oracle():
    N, S, om = _mods()
    dims, B = [13, 96, 64, 20], 100  # widths and batch off the 64 padding on purpose
    net, params, x, t = _net_and_data(S, dims, B, 5)
    for l, (w, b) in enumerate(params):
        w2, b2 = net.params(l)
        assert np.array_equal(w2, w) and np.array_equal(b2, b)
    y = net.forward(x)
    y_ref, _ = om.forward(params, x)
    assert rel_l2(y, y_ref) <= 3e-2, rel_l2(y, y_ref)
    loss = net.grads(x, t)
    g_ref, loss_ref = om.grads(params, x, t)
    assert abs(loss - loss_ref) <= 3e-2 * loss_ref
    for l in range(len(params)):
        dw, db = net.layer_grads(l)
        assert rel_l2(dw, g_ref[l][0]) <= 3e-2, (l, rel_l2(dw, g_ref[l][0]))
        assert rel_l2(db, g_ref[l][1]) <= 3e-2, (l, rel_l2(db, g_ref[l][1]))
    net.close()


def test_adam_step_against_oracle():
    """One train step: the update equals Adam applied in FP64 to the step's own gradients."""
    N, S, om = _mods()
    dims, B = [13, 64, 20], 64
    net, params, x, t = _net_and_data(S, dims, B, 6)
    lr = 1e-3
    net.train_step(x, t, lr)
    gr = [tuple(np.float64(v) for v in net.layer_grads(l)) for l in range(len(dims) - 1)]
    st = [[np.zeros_like(w, np.float64), np.zeros_like(w, np.float64), np.zeros_like(b, np.float64),
           np.zeros_like(b, np.float64)] for w, b in params]
    ref = om.adam_step(params, gr, st, 1, lr)
    for l in range(len(params)):
        w, b = net.params(l)
        # FP32 rounding of the weights (|w| < 4): 5e-7, against a step of lr = 1e-3
        assert np.max(np.abs(w - ref[l][0])) <= 5e-7
        assert np.max(np.abs(b - ref[l][1])) <= 5e-7
    net.close()


def test_training_follows_fp64_adam():
    """300 full-batch Adam steps from the same initial weights: the kernels'
    loss curve ends where the FP64 oracle's does (within 10 %, the BF16
    products' share), and both fit the smooth map."""
    N, S, om = _mods()
    rng = np.random.default_rng(7)
    A = rng.standard_normal((13, 20))
    x = rng.random((512, 13)).astype(np.float32)
    t = np.sin(x @ A).astype(np.float32)
    X, T = S.DeviceArray(x), S.DeviceArray(t)
    dims, lr, steps = [13, 128, 128, 20], 1e-3, 300
    net = S.MLP(dims, batch_max=512, seed=3)
    p = [tuple(np.float64(v) for v in net.params(l)) for l in range(len(dims) - 1)]
    first = net.train_step(X, T, lr)
    for _ in range(steps - 1):
        last = net.train_step(X, T, lr)
    net.close()
    st = [[np.zeros_like(w), np.zeros_like(w), np.zeros_like(b), np.zeros_like(b)] for w, b in p]
    for k in range(1, steps + 1):
        g, loss_ref = om.grads(p, x, t)
        if k == 1:
            assert abs(first - loss_ref) <= 2e-2 * loss_ref
        p = om.adam_step(p, g, st, k, lr)
    final_ref = om.loss(p, x, t)
    assert final_ref <= first / 5
    # `last` is the loss before the 300th update, final_ref after it: compare
    # against the oracle's loss one step earlier as well
    assert min(abs(last - final_ref), abs(last - loss_ref)) <= 0.1 * loss_ref, (first, last, loss_ref, final_ref)


def test_train_epoch_matches_steps_over_the_same_batches():
    """mlp_train_epoch (device gather, queued steps) equals train_step over the
    batches of its permutation, taken in the same order."""
    N, S, om = _mods()
    rng = np.random.default_rng(8)
    x = rng.random((300, 13)).astype(np.float32)
    t = rng.standard_normal((300, 20)).astype(np.float32)
    dims = [13, 64, 20]
    a, b = S.MLP(dims, batch_max=128, seed=9), S.MLP(dims, batch_max=128, seed=9)
    loss_a = a.train_epoch(S.DeviceArray(x), S.DeviceArray(t), 128, 77, 1e-3)
    # the library's permutation: Fisher-Yates over a SplitMix64 stream of the seed
    from mpifdtd_amd.batch import SplitMix64
    g, perm = SplitMix64(77), list(range(300))
    for i in range(299, 0, -1):
        j = g.next() % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    perm = np.array(perm)
    losses = [b.train_step(x[perm[s:s + 128]], t[perm[s:s + 128]], 1e-3) * len(perm[s:s + 128]) for s in range(0, 300, 128)]
    assert abs(loss_a - sum(losses) / 300) <= 1e-6 * abs(loss_a)
    for l in range(2):
        wa, ba = a.params(l)
        wb, bb = b.params(l)
        assert np.array_equal(wa, wb) and np.array_equal(ba, bb)
    a.close()
    b.close()


def test_configured_shape_one_training_step():
    """The surrogate's configured network, 13 -> 512 -> 512 -> 512 -> 180 at a
    batch of 1024: forward outputs and every layer's gradients against the FP64
    oracle (3e-2, the BF16 activations' share), then one Adam step equal to the
    oracle's Adam on the step's own gradients to FP32 rounding."""
    N, S, om = _mods()
    dims, B = [13, 512, 512, 512, 180], 1024
    rng = np.random.default_rng(11)
    net = S.MLP(dims, batch_max=B, seed=11)
    params = [((rng.standard_normal((dims[i + 1], dims[i])) / np.sqrt(dims[i])).astype(np.float32),
               (0.1 * rng.standard_normal(dims[i + 1])).astype(np.float32)) for i in range(len(dims) - 1)]
    for l, (w, b) in enumerate(params):
        net.set_params(l, w, b)
    x = rng.random((B, dims[0])).astype(np.float32)
    t = rng.standard_normal((B, dims[-1])).astype(np.float32)
    y = net.forward(x)
    y_ref, _ = om.forward(params, x)
    assert rel_l2(y, y_ref) <= 3e-2, rel_l2(y, y_ref)
    g_ref, loss_ref = om.grads(params, x, t)
    lr = 1e-3
    loss = net.train_step(x, t, lr)
    assert abs(loss - loss_ref) <= 3e-2 * loss_ref
    gr = []
    for l in range(len(params)):
        dw, db = net.layer_grads(l)
        assert rel_l2(dw, g_ref[l][0]) <= 3e-2, (l, rel_l2(dw, g_ref[l][0]))
        assert rel_l2(db, g_ref[l][1]) <= 3e-2, (l, rel_l2(db, g_ref[l][1]))
        gr.append((np.float64(dw), np.float64(db)))
    st = [[np.zeros_like(w, np.float64), np.zeros_like(w, np.float64), np.zeros_like(b, np.float64),
           np.zeros_like(b, np.float64)] for w, b in params]
    ref = om.adam_step(params, gr, st, 1, lr)
    for l in range(len(params)):
        w, b = net.params(l)
        assert np.max(np.abs(w - ref[l][0])) <= 5e-7
        assert np.max(np.abs(b - ref[l][1])) <= 5e-7
    net.close()
