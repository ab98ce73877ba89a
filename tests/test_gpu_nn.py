"""The network kernels (BF16 MFMA forward and backward) against the CPU
oracle, and training sanity."""

import numpy as np
import pytest

from oracle import nn_oracle as O

pytestmark = pytest.mark.gpu


def _net(b, h, w):
    from exploringsycl_amd.nn import Net
    return Net(b, h, w)


def _case(seed, b=3, h=16, w=64):
    rng = np.random.default_rng(seed)
    p = O.init_params(seed)
    p[O.num_params() - 1] = 0.05
    # non-zero biases to exercise them
    k = 0
    for o, i in O.SHAPES:
        k += o * i * 9
        p[k:k + o] = rng.uniform(-0.1, 0.1, o)
        k += o
    x = rng.standard_normal((b, h, w, 3)).astype(np.float32)
    t = rng.standard_normal((b, h, w)).astype(np.float32)
    return p, x, t


@pytest.mark.parametrize("shape", [(3, 16, 64), (2, 7, 32), (1, 64, 128), (1, 12, 512)])
def test_forward_matches_oracle(shape):
    p, x, t = _case(1, *shape)
    net = _net(*shape)
    net.set_params(p)
    got = net.forward(x)
    want, _ = O.forward(p, x, bf16_mode=True)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 2e-3, err
    assert np.allclose(net.get_params(), p)


@pytest.mark.parametrize("rows", [2, 4])
@pytest.mark.parametrize("shape", [(4, 512, 128), (1, 512, 512), (2, 256, 512)])
def test_forward_row_kernel_is_bitwise_the_one_row_kernel(shape, rows, monkeypatch):
    """The multi-row forward kernel (R output rows per block, weights kept in
    registers) against the one-row LDS kernel: same MFMA sequence per tile,
    so the outputs must be identical bit for bit."""
    p, x, _ = _case(4, *shape)
    monkeypatch.setenv("TEA_NN_FWD_ROWS", "0")
    one = _net(*shape)
    monkeypatch.setenv("TEA_NN_FWD_ROWS", str(rows))
    multi = _net(*shape)
    one.set_params(p)
    multi.set_params(p)
    a, b = one.forward(x), multi.forward(x)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("shape", [(3, 16, 64), (2, 9, 32)])
def test_gradients_match_oracle(shape):
    p, x, t = _case(2, *shape)
    net = _net(*shape)
    net.set_params(p)
    loss, g = net.loss_grad(x, t)
    loss_o, g_o = O.loss_grad(p, x, t, bf16_mode=True)
    assert abs(loss / loss_o - 1) < 2e-3
    k = 0
    for li, (o, i) in enumerate(O.SHAPES):
        n = o * i * 9 + o
        a, b = g[k:k + n], g_o[k:k + n]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < 3e-2, (li, err)
        k += n


def test_training_reduces_loss():
    """Learn a fixed linear 3x3 filter of channel 0 from random inputs."""
    rng = np.random.default_rng(3)
    b, h, w = 8, 16, 32
    net = _net(b, h, w)
    net.set_params(O.init_params(4))
    x = rng.standard_normal((64, h, w, 3)).astype(np.float32)
    k = np.array([[0.05, 0.1, 0.05], [0.1, 0.4, 0.1], [0.05, 0.1, 0.05]])
    xp = np.pad(x[..., 0], ((0, 0), (1, 1), (1, 1)))
    t = sum(k[i, j] * xp[:, i:i + h, j:j + w] for i in range(3) for j in range(3))
    t = t.astype(np.float32)
    net.load_dataset(x, t)
    first = net.train_epoch(np.arange(64), b, 3e-3)
    for _ in range(30):
        last = net.train_epoch(rng.permutation(64), b, 3e-3)
    assert last < 0.2 * first, (first, last)


# The device's Adam constants are FP32 (0.999f is 0.999 + 1.3e-8, so
# 1 - beta2 is 1e-3 (1 - 1.3e-5)); the restatement is given the same values.
ADAM_FP32 = dict(beta1=float(np.float32(0.9)), beta2=float(np.float32(0.999)),
                 eps=float(np.float32(1e-8)))
U32 = 2.0 ** -24   # unit roundoff of FP32


def _adam_step_checked(net, g, t, lr, step):
    """One device Adam step (step(g, lr)) against the float64 restatement
    applied to the device's own state before the step, within bounds of the
    FP32 roundings of each operation: m = b1 m + (1 - b1) g may cancel, so
    its bound is relative to the terms, and the parameter's to its update
    and its own spacing."""
    p0 = net.get_params().astype(np.float64)
    m0, v0, t0 = net.adam_state()
    assert t0 == t - 1
    m0, v0 = m0.astype(np.float64), v0.astype(np.float64)
    step(g, lr)
    gp = net.get_params().astype(np.float64)
    gm, gv, gt = net.adam_state()
    assert gt == t
    p, m, v = O.adam_step(p0, g, m0, v0, t, float(np.float32(lr)), **ADAM_FP32)
    b1, b2, e = ADAM_FP32["beta1"], ADAM_FP32["beta2"], ADAM_FP32["eps"]
    g64 = np.asarray(g, np.float64)
    bm = 4 * U32 * (np.abs(b1 * m0) + np.abs((1 - b1) * g64)) + 1e-44
    bv = 4 * U32 * (np.abs(b2 * v0) + np.abs((1 - b2) * g64 * g64)) + 1e-44
    assert np.all(np.abs(gm - m) <= bm)
    assert np.all(np.abs(gv - v) <= bv)
    c1, c2 = 1 - b1 ** t, 1 - b2 ** t
    den = np.sqrt(v / c2) + e
    upd = lr * (m / c1) / den
    bp = lr * (bm / c1) / den + 16 * U32 * np.abs(upd) + 2 * U32 * np.abs(p0)
    assert np.all(np.abs(gp - p) <= bp)
    return upd


def test_adam_step_matches_oracle():
    """Three Adam steps with fixed gradients (large, small, zero and near
    eps) against Kingma & Ba's algorithm in float64: master parameters,
    both moments and the step count."""
    rng = np.random.default_rng(7)
    net = _net(1, 4, 32)
    net.set_params(O.init_params(5))
    n = O.num_params()
    for t, lr in ((1, 1e-3), (2, 3e-3), (3, 1e-3)):
        g = rng.standard_normal(n) * 10.0 ** rng.uniform(-9, 1, n)
        g[:50] = 0.0
        g[50:100] = 1e-8 * np.sign(g[50:100])
        g = g.astype(np.float32)
        upd = _adam_step_checked(net, g, t, lr, net.adam_step)
        if t == 1:   # step 1 is lr sign(g) wherever |g| >> eps
            big = np.abs(g) > 1e-4
            assert np.allclose(upd[big], lr * np.sign(g[big]), rtol=1e-3)


def test_train_steps_match_oracle_adam():
    """tea_nn_train_step = the GPU gradient at the current parameters plus
    one Adam step: three steps against the restatement fed those
    gradients."""
    p0, x, t = _case(6, 2, 8, 32)
    net = _net(2, 8, 32)
    net.set_params(p0)
    for step in (1, 2, 3):
        _, g = net.loss_grad(x, t)
        _adam_step_checked(net, g, step, 1e-3, lambda g_, lr: net.train_step(x, t, lr))
