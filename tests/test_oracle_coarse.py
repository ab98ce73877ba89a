"""Pins the oracle's two-level coarse correction with Jacobi sweeps
(orc_fcg_coarse_faces / _restrict8 / _restrict4 / _sweeps / _gamma /
_prolong) against numpy restatements of what each stage means (block means,
the Galerkin product, TeaLeaf's Jacobi iteration as the training target
defines it, the energy-optimal factors as a least-squares solve) and against
exact arithmetic where the summation order cannot matter (integer data);
then the whole flexible CG with it against a direct solve, and the
stagnation fallback of the host on the oracle."""

import ctypes
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from exploringsycl_amd import learned
from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import run_deck
from tests.oracle_ctypes import ORACLE_TEALEAF
from tests.test_oracle_mpcg import _system

H = 2
pytestmark = pytest.mark.skipif(not os.path.exists(ORACLE_TEALEAF),
                                reason="oracle host not built")

_I, _L, _D = ctypes.c_int, ctypes.c_long, ctypes.c_double
_PF = ctypes.POINTER(ctypes.c_float)
_PD = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib(oracle):
    L = oracle.lib
    sig = {
        "orc_fcg_coarse_faces": [_I, _I, _I, _I, _PF, _PF, _PF],
        "orc_fcg_restrict8": [_I, _I, _I, _PF, _PF],
        "orc_fcg_restrict4": [_I, _I, _PF, _PF],
        "orc_fcg_sweeps": [_I, _I, _PF, _PF, _I, _PF, _PD, _PD],
        "orc_fcg_gamma": [_I, _PD, _I, _PD, _I, _PD, _PF, _L, _D, _D, _PF],
        "orc_fcg_prolong": [_I, _I, _I, _PF, _PF, _PF, _PF, _I, _PF, _I, _I,
                            ctypes.c_float, _PF, _I, _I, ctypes.c_float, _PF],
    }
    for name, args in sig.items():
        getattr(L, name).argtypes = args
    L.orc_fcg_prolong.restype = _D
    return L


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def pf(a):
    return a.ctypes.data_as(_PF) if a is not None else None


def pd(a):
    return a.ctypes.data_as(_PD) if a is not None else None


def _padded(rng, ny, nx, integer=False):
    x, y = nx + 2 * H, ny + 2 * H
    a = np.zeros((y, x), np.float32)
    inner = (rng.integers(-50, 50, (ny, nx)) if integer
             else rng.standard_normal((ny, nx)) * 10.0 ** rng.uniform(-3, 3, (ny, nx)))
    a[H:-H, H:-H] = inner
    return a, x, y


@pytest.mark.parametrize("integer", [True, False])
def test_restriction_is_the_block_mean(lib, integer):
    rng = np.random.default_rng(1)
    r, x, y = _padded(rng, 64, 96, integer)
    rb0 = np.zeros((8, 12), np.float32)
    lib.orc_fcg_restrict8(x, y, H, pf(r), pf(rb0))
    inner = r[H:-H, H:-H].astype(np.float64)
    want0 = learned.block_mean(inner, 8)
    rb1 = np.zeros((2, 3), np.float32)
    lib.orc_fcg_restrict4(2, 3, pf(rb0), pf(rb1))
    want1 = learned.block_mean(rb0.astype(np.float64), 4)
    if integer:   # sums of integers are exact in any order
        assert np.array_equal(rb0, want0.astype(np.float32))
        assert np.array_equal(rb1, want1.astype(np.float32))
    else:
        assert np.allclose(rb0, want0, rtol=1e-6, atol=1e-6 * np.abs(want0).max())
        assert np.allclose(rb1, want1, rtol=1e-6, atol=1e-6 * np.abs(want1).max())


@pytest.mark.parametrize("c", [8, 32])
def test_coarse_faces_are_the_galerkin_operator(lib, oracle, c):
    """The coarse faces against learned.coarse_faces, itself pinned against
    the explicit product P^T A P / c^2 (tests/test_learned.py)."""
    deck = Deck.with_random_rectangles(64, 64, 8, problem_seed(5, 7), density="log10",
                                       density_range=(-2.0, 2.0), coefficient="inverse")
    _, _, s = _system(oracle, deck)
    kx, ky = f32(s.f["kx"]), f32(s.f["ky"])
    hc = wc = 64 // c
    kc = np.zeros(2 * hc * wc, np.float32)
    lib.orc_fcg_coarse_faces(s.x, s.y, H, c, pf(kx), pf(ky), pf(kc))
    KX = s.f["kx"].reshape(s.y, s.x)[H:-H, H:s.x - H + 1]
    KY = s.f["ky"].reshape(s.y, s.x)[H:s.y - H + 1, H:-H]
    kxc, kyc = learned.coarse_faces(KX, KY, c)
    assert np.allclose(kc[:hc * wc].reshape(hc, wc), kxc[:, :-1], rtol=1e-6, atol=0)
    assert np.allclose(kc[hc * wc:].reshape(hc, wc), kyc[:-1, :], rtol=1e-6, atol=0)


def _coarse_problem(seed, hc, wc):
    """Face coefficients of a stiff coarse operator (zero on the edge) and a
    right-hand side, as the kernels store them (FP32)."""
    rng = np.random.default_rng(seed)
    kx = 10.0 ** rng.uniform(-2, 4, (hc, wc))
    ky = 10.0 ** rng.uniform(-2, 4, (hc, wc))
    kx[:, 0] = 0.0
    ky[0, :] = 0.0
    kc = f32(np.concatenate([kx.ravel(), ky.ravel()]))
    rbar = f32(rng.standard_normal((hc, wc)))
    return kc, rbar


def _numpy_operator(kc, hc, wc):
    """Padded (kx, ky) layouts of learned.jacobi_sweeps and T as a dense
    matrix, from the stored faces."""
    n = hc * wc
    kl = kc[:n].reshape(hc, wc).astype(np.float64)
    kb = kc[n:].reshape(hc, wc).astype(np.float64)
    # TeaLeaf's layout: kx[j, k] is the face left of cell k (the right face
    # of the last cell, in the halo, stays zero), ky[j, k] the face below.
    KX = np.zeros((hc + 2 * H, wc + 2 * H))
    KY = np.zeros((hc + 2 * H, wc + 2 * H))
    KX[H:-H, H:H + wc] = kl
    KY[H:H + hc, H:-H] = kb
    T = np.zeros((n, n))
    for Y in range(hc):
        for X in range(wc):
            i = Y * wc + X
            T[i, i] += 1.0
            for (j, k) in ((i - 1, kl[Y, X]) if X > 0 else (None, 0.0),
                           (i - wc, kb[Y, X]) if Y > 0 else (None, 0.0)):
                if j is not None:
                    T[i, i] += k
                    T[j, j] += k
                    T[i, j] -= k
                    T[j, i] -= k
    return KX, KY, T


@pytest.mark.parametrize("sweeps", [1, 2, 6])
@pytest.mark.parametrize("hc,wc", [(16, 16), (20, 36), (8, 12)])
def test_sweeps_are_teleafs_jacobi_on_the_coarse_operator(lib, sweeps, hc, wc):
    kc, rbar = _coarse_problem(sweeps * 100 + hc, hc, wc)
    ntiles = ((wc + 15) // 16) * ((hc + 15) // 16)
    y = np.zeros(hc * wc, np.float32)
    part = np.zeros(2 * ntiles)
    cross = np.zeros((hc // 4) * (wc // 4))
    lib.orc_fcg_sweeps(hc, wc, pf(kc), pf(rbar), sweeps, pf(y), pd(part), pd(cross))
    KX, KY, T = _numpy_operator(kc, hc, wc)
    rhs = np.zeros((hc + 2 * H, wc + 2 * H))
    rhs[H:-H, H:-H] = rbar
    want = learned.jacobi_sweeps(rhs, KX, KY, sweeps=sweeps)[H:-H, H:-H].ravel()
    assert np.allclose(y, want, rtol=2e-7, atol=1e-7 * np.abs(want).max())
    if sweeps == 1:   # one sweep from zero is rbar / diag(T)
        assert np.allclose(y, rbar.ravel() / np.diag(T), rtol=1e-7)
    yd = y.astype(np.float64)
    Ty = T @ yd
    assert np.isclose(part[:ntiles].sum(), yd @ rbar.ravel().astype(np.float64), rtol=1e-12)
    assert np.isclose(part[ntiles:].sum(), yd @ Ty, rtol=1e-10)
    blocks = Ty.reshape(hc // 4, 4, wc // 4, 4).sum(axis=(1, 3)) if hc % 4 == 0 else None
    if blocks is not None:
        assert np.allclose(cross, blocks.ravel(), rtol=1e-9, atol=1e-9 * np.abs(Ty).max())


def test_many_sweeps_solve_the_coarse_system(lib):
    """Jacobi converges on T (strictly diagonally dominant): many sweeps
    reach T^-1 rbar."""
    hc, wc = 8, 8
    kc, rbar = _coarse_problem(5, hc, wc)
    kc = f32(np.minimum(kc, 3.0))  # a contraction factor well below 1
    y = np.zeros(hc * wc, np.float32)
    part = np.zeros(2)
    # 8 sweeps per call at most on the device; the oracle takes any count.
    lib.orc_fcg_sweeps(hc, wc, pf(kc), pf(rbar), 400, pf(y), pd(part), None)
    _, _, T = _numpy_operator(kc, hc, wc)
    want = np.linalg.solve(T, rbar.ravel().astype(np.float64))
    assert np.allclose(y, want, rtol=1e-6, atol=1e-6 * np.abs(want).max())


def test_gamma_is_the_energy_optimal_combination(lib):
    """g solves min_g |e - g0 d0 - g1 d1|_A: the 2 x 2 Gram system of the
    directions, assembled here from dense vectors."""
    rng = np.random.default_rng(3)
    n0, n1, nx = 5, 3, 40
    part0 = rng.standard_normal(2 * n0)
    part0[n0:] = np.abs(part0[n0:]) + 1.0
    part1 = rng.standard_normal(2 * n1)
    part1[n1:] = np.abs(part1[n1:]) + 1.0
    cross = rng.standard_normal(nx) * 0.1
    y1 = f32(rng.standard_normal(nx))
    g = np.zeros(2, np.float32)
    c0sq, c1sq = 64.0, 1024.0
    lib.orc_fcg_gamma(2, pd(part0), n0, pd(part1), n1, pd(cross), pf(y1), nx, c0sq,
                      c1sq, pf(g))
    a = np.array([c0sq * part0[:n0].sum(), c1sq * part1[:n1].sum()])
    b12 = c0sq * cross @ y1.astype(np.float64)
    B = np.array([[c0sq * part0[n0:].sum(), b12], [b12, c1sq * part1[n1:].sum()]])
    assert np.allclose(g, np.linalg.solve(B, a), rtol=1e-6)
    lib.orc_fcg_gamma(1, pd(part0), n0, None, 0, None, None, 0, c0sq, c1sq, pf(g))
    assert np.isclose(g[0], part0[:n0].sum() / part0[n0:].sum(), rtol=1e-6) and g[1] == 0
    zero = np.zeros(2 * n0)   # y = 0: no coarse direction, g = 0 (not NaN)
    lib.orc_fcg_gamma(1, pd(zero), n0, None, 0, None, None, 0, c0sq, c1sq, pf(g))
    assert g[0] == 0.0


def test_prolongation_adds_the_scaled_coarse_values(lib):
    rng = np.random.default_rng(4)
    ny, nx = 64, 64
    x, y = nx + 2 * H, ny + 2 * H
    r = f32(rng.standard_normal(x * y))
    w = f32(rng.standard_normal(x * y))
    kx = f32(rng.uniform(0, 5, x * y))
    ky = f32(rng.uniform(0, 5, x * y))
    y0 = f32(rng.standard_normal(64))
    y1 = f32(rng.standard_normal(4))
    z = np.zeros(x * y, np.float32)
    zw = lib.orc_fcg_prolong(x, y, H, pf(r), pf(kx), pf(ky), pf(w), 2, pf(y0), 8, 8,
                             0.5, pf(y1), 2, 32, -0.25, pf(z))
    K = kx.reshape(y, x).astype(np.float64)
    Ky = ky.reshape(y, x).astype(np.float64)
    d = (1.0 + (K[H:-H, H + 1:x - H + 1] + K[H:-H, H:-H])) + (Ky[H + 1:y - H + 1, H:-H] + Ky[H:-H, H:-H])
    jj, kk = np.mgrid[0:ny, 0:nx]
    want = (r.reshape(y, x)[H:-H, H:-H] / d + 0.5 * y0.reshape(8, 8)[jj // 8, kk // 8]
            - 0.25 * y1.reshape(2, 2)[jj // 32, kk // 32])
    Z = z.reshape(y, x)
    assert np.allclose(Z[H:-H, H:-H], want, rtol=1e-6, atol=1e-6)
    outside = np.ones((y, x), bool)
    outside[H:-H, H:-H] = False
    assert not Z[outside].any()
    assert np.isclose(zw, (Z[H:-H, H:-H].astype(np.float64) * w.reshape(y, x)[H:-H, H:-H]).sum(),
                      rtol=1e-12)


def _fcg_deck(n, seed, extra, count=24):
    return Deck.with_random_rectangles(n, n, count, problem_seed(5, seed), density="log10",
                                       density_range=(-2.0, 2.0), end_step=1,
                                       coefficient="inverse", extent=10.0 * n / 4096,
                                       extra=["rel_eps=1e-10"] + extra)


@pytest.mark.parametrize("n,sweeps,levels", [(64, 6, 2), (96, 6, 2), (128, 1, 2), (64, 3, 1)])
def test_two_level_fcg_reaches_the_direct_solution(tmp_path, oracle, n, sweeps, levels):
    deck = _fcg_deck(n, n, ["use_fcg", f"fcg_coarse_sweeps={sweeps}", f"fcg_levels={levels}"])
    rec, _ = run_deck(deck, binary=ORACLE_TEALEAF, workdir=str(tmp_path),
                      extra_args=["--field-out", "u.bin"])
    assert rec["coarse_solve"] == f"jacobi sweeps {sweeps}"
    st = rec["steps"][0]
    assert st["rel_residual"] < 1e-10 and st["fallback_at"] == -1
    u = np.fromfile(tmp_path / "u.bin")
    A, u0, _ = _system(oracle, deck)
    exact = spla.spsolve(A.tocsc(), u0)
    assert np.linalg.norm(u - exact) <= 1e-10 * np.linalg.norm(u0)


def test_two_level_fcg_needs_fewer_iterations_than_jacobi_pcg(tmp_path):
    def run(extra, name):
        rec, _ = run_deck(_fcg_deck(256, 3000, extra, count=64), binary=ORACLE_TEALEAF,
                          workdir=str(tmp_path / name))
        return rec["steps"][0]["iterations"]
    two = run(["use_fcg", "fcg_coarse_sweeps=6"], "two")
    jac = run(["use_jacobi_pcg"], "jac")
    assert two * 3 < jac, (two, jac)


def test_stagnation_falls_back_to_jacobi_pcg(tmp_path):
    """A stall rule no method can meet (|r| down 1000x every 20 iterations)
    must fire at the end of the first window; the solve restarts (beta = 0)
    as Jacobi-PCG and still reaches 1e-10, at the Jacobi-PCG answer."""
    deck = _fcg_deck(96, 11, ["use_fcg", "fcg_coarse_sweeps=6", "stall_window=20",
                              "stall_factor=1000"])
    rec, _ = run_deck(deck, binary=ORACLE_TEALEAF, workdir=str(tmp_path / "fcg"))
    st = rec["steps"][0]
    assert st["fallback_at"] == 20
    assert st["fallback_beta"] == 0.0
    assert st["rel_residual"] < 1e-10
    jp, _ = run_deck(_fcg_deck(96, 11, ["use_jacobi_pcg"]), binary=ORACLE_TEALEAF,
                     workdir=str(tmp_path / "jp"))
    assert abs(rec["summaries"][-1]["temp"] / jp["summaries"][-1]["temp"] - 1) < 1e-8
    # The literal rule of the specification (2x over 20 iterations) also
    # fires: no method reaches 1e-10 at that rate on this workload.
    lit, _ = run_deck(_fcg_deck(96, 11, ["use_fcg", "fcg_coarse_sweeps=6", "stall_window=20",
                                         "stall_factor=2"]),
                      binary=ORACLE_TEALEAF, workdir=str(tmp_path / "lit"))
    assert lit["steps"][0]["fallback_at"] > 0
    assert lit["steps"][0]["rel_residual"] < 1e-10
