"""Pins the oracle's mixed-precision CG (orc_mpcg_*, orc_dot32) on its own,
without the GPU: the whole solve against a sparse direct solve of the same
system, and the FP64 true-residual refresh against numpy's u0 - A u.

The operator is assembled in scipy from the oracle's face coefficients and
checked against the oracle's own (pinned) stencil on random vectors before
it is used, so a wrong assembly cannot make both sides agree."""

import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import run_deck
from tests.oracle_ctypes import ORACLE_TEALEAF, OracleSolver

H = 2
pytestmark = pytest.mark.skipif(not os.path.exists(ORACLE_TEALEAF),
                                reason="oracle host not built")


def _deck(n, seed, extra):
    if seed is None:   # the tea.in states, K = density
        return Deck(n, n, end_step=1, extra=extra)
    return Deck.with_random_rectangles(n, n, 12, problem_seed(3, seed), density="log10",
                                       density_range=(-2.0, 2.0), end_step=1,
                                       coefficient="inverse", extra=extra)


def _system(oracle, deck):
    """(A as scipy CSR over the interior, u0 interior, solver) for the first
    timestep of ``deck``, built by the oracle's cg_init."""
    coef = 2 if deck.coefficient == "inverse" else 1
    s = OracleSolver(oracle, deck.x_cells, deck.y_cells, deck.states, coefficient=coef,
                     h=H, dt=deck.dt, extent=deck.extent)
    oracle.cg_init(s.x, s.y, H, coef, s.rx, s.ry, s.f)
    nx, ny = deck.x_cells, deck.y_cells
    kx = s.f["kx"].reshape(s.y, s.x)
    ky = s.f["ky"].reshape(s.y, s.x)
    idx = np.arange(nx * ny).reshape(ny, nx)
    rows, cols, vals = [], [], []
    diag = np.ones((ny, nx))
    for j in range(ny):
        for k in range(nx):
            J, K = j + H, k + H
            # Face coefficients of the cell's four faces; a face on the
            # domain edge couples to the mirrored halo cell, i.e. to itself.
            for (jj, kk, c) in ((j, k + 1, kx[J, K + 1]), (j, k - 1, kx[J, K]),
                                (j + 1, k, ky[J + 1, K]), (j - 1, k, ky[J, K])):
                diag[j, k] += c
                if 0 <= jj < ny and 0 <= kk < nx:
                    rows.append(idx[j, k]); cols.append(idx[jj, kk]); vals.append(-c)
                else:
                    diag[j, k] -= c
    A = sp.csr_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny)) + sp.diags(diag.ravel())
    u0 = s.f["u"].reshape(s.y, s.x)[H:s.y - H, H:s.x - H].ravel().copy()
    # The assembly against the oracle's own stencil (cg_calc_w on a field
    # whose halo is reflected), on random vectors.
    rng = np.random.default_rng(0)
    for _ in range(3):
        v = rng.standard_normal((ny, nx))
        p = np.zeros(s.x * s.y)
        p.reshape(s.y, s.x)[H:s.y - H, H:s.x - H] = v
        oracle.halo_field(s.x, s.y, H, 1, p)
        g = dict(s.f, p=p, w=np.zeros(s.x * s.y))
        oracle.cg_calc_w(s.x, s.y, H, g)
        Av = g["w"].reshape(s.y, s.x)[H:s.y - H, H:s.x - H].ravel()
        assert np.allclose(A @ v.ravel(), Av, rtol=1e-13, atol=1e-13 * np.abs(Av).max())
    return A, u0, s


@pytest.mark.parametrize("n,seed", [(10, None), (64, None), (64, 1), (48, 2)])
def test_mpcg_reaches_the_direct_solution(tmp_path, oracle, n, seed):
    """The oracle's mixed-precision CG, run by this project's host, stops at
    |b - A u| < 1e-10 |b|; since A = I + (SPD), |A^-1| <= 1 and the final u
    is within 1e-10 |b| of the exact solution."""
    deck = _deck(n, seed, ["use_mpcg", "rel_eps=1e-10"])
    rec, _ = run_deck(deck, binary=ORACLE_TEALEAF, workdir=str(tmp_path),
                      extra_args=["--field-out", "u.bin"])
    assert rec["solver"] == "MPCG"
    step = rec["steps"][0]
    assert step["rel_residual"] < 1e-10
    u = np.fromfile(tmp_path / "u.bin", dtype=np.float64)
    A, u0, _ = _system(oracle, deck)
    exact = spla.spsolve(A.tocsc(), u0)
    assert np.linalg.norm(A @ exact - u0) < 1e-13 * np.linalg.norm(u0)
    # The stop test itself, recomputed: the true relative residual of u.
    rel = np.linalg.norm(u0 - A @ u) / np.linalg.norm(u0)
    assert rel < 1e-10 and np.isclose(rel, step["rel_residual"], rtol=1e-3)
    assert np.linalg.norm(u - exact) <= 1e-10 * np.linalg.norm(u0)
    assert np.abs(u - exact).max() <= 1e-9 * np.abs(exact).max()


def _bind_refresh(oracle):
    lib = oracle.lib
    f = lib.orc_mpcg_refresh
    D, I = ctypes.c_double, ctypes.c_int
    PD, PF = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
    f.argtypes = [I, I, I, D, I, PD, PF, PD, PD, PD, PF, PF, PD]
    f.restype = None
    return f


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


@pytest.mark.parametrize("apply_u", [0, 1])
def test_mpcg_refresh_is_the_fp64_true_residual(oracle, apply_u):
    """orc_mpcg_refresh: (optionally u += alpha p first, with the halo
    reflected) r = u0 - A u in FP64 rounded once to FP32, with the sums
    r32.r32, r.r (FP64) and r32.w against numpy."""
    deck = _deck(40, 4, [])
    A, _, s = _system(oracle, deck)
    x, y = s.x, s.y
    rng = np.random.default_rng(3)
    u = np.zeros(x * y)
    u.reshape(y, x)[H:y - H, H:x - H] = rng.uniform(0.5, 2.0, (y - 2 * H, x - 2 * H))
    oracle.halo_field(x, y, H, 1, u)
    u0 = rng.uniform(0.5, 2.0, x * y)
    p = rng.standard_normal(x * y).astype(np.float32)
    w = rng.standard_normal(x * y).astype(np.float32)
    r32 = np.zeros(x * y, dtype=np.float32)
    alpha = 0.37
    u_before = u.copy()
    out = np.zeros(3)
    _bind_refresh(oracle)(x, y, H, alpha, apply_u, _ptr(u, ctypes.c_double),
                          _ptr(p, ctypes.c_float), _ptr(u0, ctypes.c_double),
                          _ptr(s.f["kx"], ctypes.c_double), _ptr(s.f["ky"], ctypes.c_double),
                          _ptr(r32, ctypes.c_float), _ptr(w, ctypes.c_float),
                          _ptr(out, ctypes.c_double))
    inner = (slice(H, y - H), slice(H, x - H))
    want_u = u_before.reshape(y, x)[inner].copy()
    if apply_u:
        want_u = want_u + alpha * p.reshape(y, x)[inner].astype(np.float64)
    assert np.array_equal(u.reshape(y, x)[inner], want_u)
    if apply_u:   # the depth-1 ring is the reflection of the updated interior
        U = u.reshape(y, x)
        assert np.array_equal(U[H:y - H, H - 1], U[H:y - H, H])
        assert np.array_equal(U[H - 1, H:x - H], U[H, H:x - H])
    rt = u0.reshape(y, x)[inner].ravel() - A @ want_u.ravel()
    got = r32.reshape(y, x)[inner].ravel()
    # FP64 residuals that differ in the last bits round to the same FP32
    # value or to a neighbour.
    assert np.all(np.abs(got - rt.astype(np.float32)) <= np.spacing(np.abs(got)))
    outside = np.ones((y, x), bool)
    outside[inner] = False
    assert not r32.reshape(y, x)[outside].any()
    g64 = got.astype(np.float64)
    assert np.isclose(out[0], np.dot(g64, g64), rtol=1e-12)
    assert np.isclose(out[1], np.dot(rt, rt), rtol=1e-12)
    wi = w.reshape(y, x)[inner].ravel().astype(np.float64)
    assert np.isclose(out[2], np.dot(g64, wi), rtol=1e-10, atol=1e-12 * np.abs(g64 * wi).sum())
