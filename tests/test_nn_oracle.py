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
