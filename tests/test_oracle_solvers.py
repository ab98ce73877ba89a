"""Pins the oracle's Jacobi, Chebyshev and PPCG kernels (TeaLeaf/c_kernels/
sycl/jacobi.cpp, cheby.cpp, ppcg.cpp) and this project's drivers for them
with fixed points, invariants, the Chebyshev error bound, the Ritz-value
bracket of the eigenvalue estimate, and the reference's known answers run
with each solver.  The dense operator comes from the oracle's own 5-point
operator, itself pinned in test_oracle_operator.py."""

import math

import numpy as np
import pytest

from exploringsycl_amd.deck import Deck, TEA_IN_STATES
from exploringsycl_amd.runner import run_deck
from tests.oracle_ctypes import ORACLE_TEALEAF, OracleSolver
from tests.test_oracle_operator import H, _apply_xy

ENV = {"OMP_NUM_THREADS": "8"}


def _setup(oracle, n, coef=1, states=TEA_IN_STATES):
    s = OracleSolver(oracle, n, n, states, coefficient=coef)
    return s


def _inner(s, a):
    return a.reshape(s.y, s.x)[H:s.y - H, H:s.x - H]


def _dense(oracle, s, n):
    A = np.zeros((n * n, n * n))
    for j in range(n * n):
        e = np.zeros(n * n)
        e[j] = 1.0
        A[:, j] = _apply_xy(oracle, s.x, s.y, s.f, e.reshape(n, n)).ravel()
    return A


def _reflect(oracle, s, name):
    oracle.halo_field(s.x, s.y, H, 1, s.f[name])


def _cheby_coefficients(lmin, lmax, count):
    """alpha_k = rho_(k+1) rho_k, beta_k = 2 rho_(k+1) / delta with
    rho_k = T_k(sigma) / T_(k+1)(sigma) from T_k(sigma) = cosh(k acosh sigma)
    (sigma > 1), i.e. the closed form of the reference's recurrence."""
    theta, delta = (lmax + lmin) / 2, (lmax - lmin) / 2
    sigma = theta / delta
    t = [math.cosh(k * math.acosh(sigma)) for k in range(count + 2)]
    rho = [t[k] / t[k + 1] for k in range(count + 1)]
    alphas = [rho[k + 1] * rho[k] for k in range(count)]
    betas = [2 * rho[k + 1] / delta for k in range(count)]
    return theta, sigma, alphas, betas


# ---- Jacobi ---------------------------------------------------------------------


@pytest.mark.parametrize("coef", [1, 2])
def test_jacobi_init_coefficients_are_cg_init_coefficients(oracle, coef):
    """jacobi_init builds the same face coefficients as cg_init (pinned by the
    operator tests) and u0 = u = energy * density inside the outer ring."""
    s = _setup(oracle, 14, coef)
    g = {k: v.copy() for k, v in s.f.items()}
    oracle.cg_init(s.x, s.y, H, coef, s.rx, s.ry, s.f)
    oracle.jacobi_init(s.x, s.y, H, coef, s.rx, s.ry, g)
    assert np.array_equal(s.f["kx"], g["kx"])
    assert np.array_equal(s.f["ky"], g["ky"])
    ring = (slice(1, s.y - 1), slice(1, s.x - 1))
    want = (g["energy"] * g["density"]).reshape(s.y, s.x)[ring]
    assert np.array_equal(g["u0"].reshape(s.y, s.x)[ring], want)
    assert np.array_equal(g["u"].reshape(s.y, s.x)[ring], want)


@pytest.mark.parametrize("coef", [1, 2])
def test_jacobi_fixed_point_is_the_dense_solution(oracle, coef):
    n = 12
    s = _setup(oracle, n, coef)
    oracle.jacobi_init(s.x, s.y, H, coef, s.rx, s.ry, s.f)
    A = _dense(oracle, s, n)
    ustar = np.linalg.solve(A, _inner(s, s.f["u0"]).ravel())
    _inner(s, s.f["u"])[:] = ustar.reshape(n, n)
    _reflect(oracle, s, "u")
    err = oracle.jacobi_iterate(s.x, s.y, H, s.f)
    assert np.allclose(_inner(s, s.f["u"]).ravel(), ustar, rtol=1e-13, atol=0)
    assert err < 1e-12 * np.abs(ustar).sum()


def test_jacobi_contracts_to_the_solution(oracle):
    """Rows are strictly diagonally dominant (1 + sum k > sum k), so each
    sweep shrinks the max-norm error; r is the previous iterate on every
    cell; the error returned is the L1 norm of the update."""
    n = 12
    s = _setup(oracle, n, 2)
    oracle.jacobi_init(s.x, s.y, H, 2, s.rx, s.ry, s.f)
    A = _dense(oracle, s, n)
    ustar = np.linalg.solve(A, _inner(s, s.f["u0"]).ravel())
    _reflect(oracle, s, "u")
    prev = np.abs(_inner(s, s.f["u"]).ravel() - ustar).max()
    for _ in range(60):
        before = s.f["u"].copy()
        err = oracle.jacobi_iterate(s.x, s.y, H, s.f)
        assert np.array_equal(s.f["r"], before)
        step = np.abs(_inner(s, s.f["u"]) - _inner(s, before)).sum()
        assert np.isclose(err, step, rtol=1e-12, atol=0)
        _reflect(oracle, s, "u")
        cur = np.abs(_inner(s, s.f["u"]).ravel() - ustar).max()
        assert cur < prev or cur < 1e-13 * np.abs(ustar).max()
        prev = cur
    assert prev < 1e-6 * np.abs(ustar).max()


# ---- Chebyshev ------------------------------------------------------------------


@pytest.mark.parametrize("coef", [1, 2])
def test_chebyshev_meets_its_error_bound(oracle, coef):
    """With the exact extreme eigenvalues, cheby_init + (alpha, beta) = (1, 0)
    (x1 = x0 + r0/theta) and then the recurrence's coefficients is the
    Chebyshev semi-iteration, whose error obeys |e_k| <= |e_0| / T_k(sigma)
    for every k; w and r hold A u and b - A u of the iterate before the step."""
    n = 12
    s = _setup(oracle, n, coef)
    oracle.cg_init(s.x, s.y, H, coef, s.rx, s.ry, s.f)
    oracle.copy_u(s.x, s.y, H, s.f)  # u0 = b
    A = _dense(oracle, s, n)
    lam = np.linalg.eigvalsh(A)
    K = 25
    theta, sigma, alphas, betas = _cheby_coefficients(lam[0], lam[-1], K)
    b = _inner(s, s.f["u0"]).ravel().copy()
    ustar = np.linalg.solve(A, b)
    _inner(s, s.f["u"])[:] = 0.0
    _reflect(oracle, s, "u")
    e0 = np.linalg.norm(ustar)
    oracle.cheby_init(s.x, s.y, H, theta, s.f)
    steps = [(1.0, 0.0)] + list(zip(alphas, betas))[:K - 1]
    for k, (a, bt) in enumerate(steps, start=1):
        u_prev = _inner(s, s.f["u"]).ravel().copy()
        oracle.cheby_iterate(s.x, s.y, H, a, bt, s.f)
        _reflect(oracle, s, "u")
        assert np.allclose(_inner(s, s.f["w"]).ravel(), A @ u_prev, rtol=1e-12,
                           atol=1e-12 * np.abs(b).max())
        assert np.allclose(_inner(s, s.f["r"]).ravel(), b - A @ u_prev, rtol=1e-10,
                           atol=1e-12 * np.abs(b).max())
        ek = np.linalg.norm(_inner(s, s.f["u"]).ravel() - ustar)
        bound = e0 / math.cosh(k * math.acosh(sigma))
        assert ek <= bound * (1 + 1e-6) + 1e-12 * e0, (k, ek, bound)


# ---- PPCG inner steps -------------------------------------------------------------


@pytest.mark.parametrize("coef", [1, 2])
def test_ppcg_inner_steps_are_a_chebyshev_solve_of_the_residual(oracle, coef):
    """ppcg_init + k inner steps accumulate z = u - u_start, the k-step
    Chebyshev approximation of A^-1 r0: r stays r0 - A z, and
    |r_k| <= |r0| / T_k(sigma)."""
    n = 12
    s = _setup(oracle, n, coef)
    oracle.cg_init(s.x, s.y, H, coef, s.rx, s.ry, s.f)  # r0 = b - A u
    A = _dense(oracle, s, n)
    lam = np.linalg.eigvalsh(A)
    K = 12
    theta, sigma, alphas, betas = _cheby_coefficients(lam[0], lam[-1], K)
    r0 = _inner(s, s.f["r"]).ravel().copy()
    u_start = _inner(s, s.f["u"]).ravel().copy()
    oracle.ppcg_init(s.x, s.y, H, theta, s.f)
    for k in range(K):
        _reflect(oracle, s, "sd")
        oracle.ppcg_inner_iteration(s.x, s.y, H, alphas[k], betas[k], s.f)
        z = _inner(s, s.f["u"]).ravel() - u_start
        r = _inner(s, s.f["r"]).ravel()
        assert np.allclose(r, r0 - A @ z, rtol=0, atol=1e-11 * np.abs(r0).max())
        bound = np.linalg.norm(r0) / math.cosh((k + 1) * math.acosh(sigma))
        assert np.linalg.norm(r) <= bound * (1 + 1e-6), (k, np.linalg.norm(r), bound)


# ---- drivers ------------------------------------------------------------------------


@pytest.mark.parametrize("solver", ["use_chebyshev", "use_ppcg"])
def test_eigenvalue_estimate_brackets_inside_the_spectrum(oracle, solver):
    """The bounds printed by the host are 0.95 x the smallest and 1.05 x the
    largest Ritz value of the CG steps' Lanczos matrix; Ritz values lie
    inside [lambda_min(A), lambda_max(A)], and the largest is within 15% of
    lambda_max after the 16 CG steps taken before the switch."""
    n = 48
    rec, _ = run_deck(Deck(n, n, end_step=1, extra=[solver, "presteps 15"]),
                      binary=ORACLE_TEALEAF, env=ENV)
    step = rec["steps"][0]
    assert step["solver_iterations"] > 0
    s = _setup(oracle, n)
    oracle.cg_init(s.x, s.y, H, 1, s.rx, s.ry, s.f)
    lam = np.linalg.eigvalsh(_dense(oracle, s, n))
    ritz_min, ritz_max = step["eigmin"] / 0.95, step["eigmax"] / 1.05
    assert lam[0] * (1 - 1e-12) <= ritz_min <= ritz_max <= lam[-1] * (1 + 1e-12)
    assert ritz_max > 0.85 * lam[-1]


@pytest.mark.parametrize("solver", ["use_jacobi", "use_chebyshev", "use_ppcg"])
@pytest.mark.parametrize("cells,steps", [(10, 10), (64, 1)])
def test_known_answers_with_each_solver(solver, cells, steps):
    """TeaLeaf/tea.problems: every solver converges to the published
    temperature: the host's own 1e-5 check, and 1e-6 relative (these
    solvers stop when the squared residual, not its root, is below eps,
    as the reference's drivers do, so they stop earlier than CG)."""
    rec, _ = run_deck(Deck(cells, cells, end_step=steps, extra=[solver]),
                      binary=ORACLE_TEALEAF, env=ENV)
    assert rec["check_found"] and rec["check_passed"]
    temp = rec["summaries"][-1]["temp"]
    assert abs(temp / rec["check_value"] - 1.0) < 1e-6, (temp, rec["check_value"])
