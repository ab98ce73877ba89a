"""Pins the CPU oracle of the preconditioner network: its analytic gradient
against central finite differences (float64, no BF16 rounding), its
parameter count, and the BF16 rounding helper."""

import numpy as np

from oracle import nn_oracle as O


def test_parameter_count():
    assert O.num_params() == 38177


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1.
    assert O.bf16(np.float32(1 + 2 ** -8)) == np.float32(1.0)
    assert O.bf16(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)
    assert O.bf16(np.float32(-2.5)) == np.float32(-2.5)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = O.init_params(1).astype(np.float64)
    p[-1] = 0.3
    x = rng.standard_normal((2, 6, 5, 3))
    t = rng.standard_normal((2, 6, 5))
    _, g = O.loss_grad(p, x, t, bf16_mode=False)
    scale = np.abs(g).max()
    for k in rng.choice(p.size, 60, replace=False):
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.loss_grad(pp, x, t, False)[0] - O.loss_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * scale + 1e-4 * abs(g[k]), (k, fd, g[k])


def test_transpose_is_adjoint():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((1, 5, 7, 4))
    d = rng.standard_normal((1, 5, 7, 3))
    w = rng.standard_normal((3, 4, 3, 3))
    lhs = np.sum(O.conv3x3(a, w) * d)
    rhs = np.sum(a * O.conv3x3_transpose(d, w))
    assert np.isclose(lhs, rhs, rtol=1e-12)


def test_adam_first_step_is_a_signed_step_of_lr():
    # With bias correction, step 1 has m_hat = g and v_hat = g^2, so every
    # parameter moves by lr * g / (|g| + eps): lr * sign(g) for |g| >> eps,
    # nothing for g = 0, and lr / 2 at |g| = eps.
    g = np.array([3.0, -0.25, 0.0, 1e-8, -4e-3])
    p, m, v = O.adam_step(np.zeros(5), g, np.zeros(5), np.zeros(5), 1, 0.01)
    assert np.allclose(p[[0, 1, 4]], [-0.01, 0.01, 0.01], rtol=1e-5)
    assert p[2] == 0.0
    assert np.isclose(p[3], -0.005, rtol=1e-6)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_adam_constant_gradient_moves_by_lr_every_step():
    # With a constant gradient the bias-corrected moments equal g and g^2 at
    # every step (m_t = (1 - b1^t) g exactly), so each step is lr * sign(g);
    # a missing bias correction would shrink the early steps.
    p, m, v = np.array([1.0, -2.0]), np.zeros(2), np.zeros(2)
    g = np.array([0.5, -7.0])
    for t in range(1, 6):
        before = p
        p, m, v = O.adam_step(p, g, m, v, t, 1e-3)
        assert np.allclose(before - p, 1e-3 * np.sign(g), rtol=1e-7)
        assert np.allclose(m, (1 - 0.9 ** t) * g, rtol=1e-12)
        assert np.allclose(v, (1 - 0.999 ** t) * g * g, rtol=1e-12)


def test_adam_second_step_closed_form():
    # Step 2 after g1 then g2: m_hat = (b1 g1 + g2)/(1 + b1),
    # v_hat = (b2 g1^2 + g2^2)/(1 + b2); this pins both betas.
    g1, g2 = np.array([1.0, 0.2]), np.array([-0.5, 0.6])
    p, m, v = O.adam_step(np.zeros(2), g1, np.zeros(2), np.zeros(2), 1, 0.1)
    p2, _, _ = O.adam_step(p, g2, m, v, 2, 0.1)
    mh = (0.9 * g1 + g2) / 1.9
    vh = (0.999 * g1 ** 2 + g2 ** 2) / 1.999
    assert np.allclose(p - p2, 0.1 * mh / (np.sqrt(vh) + 1e-8), rtol=1e-12)
