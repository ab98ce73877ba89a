"""Pins the oracle's stencil, CG, halos and field summary with closed forms,
invariants and brute force (no retyped formulas)."""

import numpy as np
import pytest

from exploringsycl_amd.deck import TEA_IN_STATES
from tests.oracle_ctypes import OracleSolver

H = 2


def _solver(oracle, n, coef=1, states=TEA_IN_STATES):
    return OracleSolver(oracle, n, n, states, coefficient=coef)


def _interior(s, a):
    return a.reshape(s.y, s.x)[H:s.y - H, H:s.x - H]


def _apply_xy(oracle, x, y, f, v):
    """A v on the interior of an x*y chunk whose kx, ky are in f, with v's
    reflective halo filled first."""
    p = np.zeros(x * y)
    p.reshape(y, x)[H:y - H, H:x - H] = v
    oracle.halo_field(x, y, H, 1, p)
    g = dict(f, p=p, w=np.zeros(x * y))
    oracle.cg_calc_w(x, y, H, g)
    return g["w"].reshape(y, x)[H:y - H, H:x - H].copy()


def _apply(oracle, s, v):
    return _apply_xy(oracle, s.x, s.y, s.f, v)


@pytest.mark.parametrize("coef", [1, 2])
def test_operator_symmetric_rowsum_one(oracle, coef):
    s = _solver(oracle, 20, coef)
    oracle.cg_init(s.x, s.y, H, coef, s.rx, s.ry, s.f)  # builds kx, ky
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal((2, 20, 20))
    Aa, Ab = _apply(oracle, s, a), _apply(oracle, s, b)
    assert np.isclose((b * Aa).sum(), (a * Ab).sum(), rtol=1e-12)
    ones = _apply(oracle, s, np.ones((20, 20)))
    assert np.allclose(ones, 1.0, rtol=0, atol=1e-9)
    # Positive definite: a'Aa > a'a (A = I + a Laplacian).
    assert (a * Aa).sum() > (a * a).sum()


def test_cg_matches_dense_solve(oracle):
    """One timestep of CG on 10x10 against numpy's dense solve of A u = u0,
    A assembled column by column from the oracle's own operator."""
    s = _solver(oracle, 10)
    x, y, f = s.x, s.y, s.f
    rro = oracle.cg_init(x, y, H, 1, s.rx, s.ry, f)
    for name in ("u", "p"):
        oracle.halo_field(x, y, H, 1, f[name])
    oracle.copy_u(x, y, H, f)
    b = _interior(s, f["u0"]).ravel().copy()
    for _ in range(200):
        pw = oracle.cg_calc_w(x, y, H, f)
        alpha = rro / pw
        rrn = oracle.cg_calc_ur(x, y, H, alpha, f)
        oracle.cg_calc_p(x, y, H, rrn / rro, f)
        for name in ("u", "p"):
            oracle.halo_field(x, y, H, 1, f[name])
        rro = rrn
        if np.sqrt(abs(rrn)) < 1e-15:
            break
    u = _interior(s, f["u"]).ravel().copy()
    A = np.zeros((100, 100))
    for j in range(100):
        e = np.zeros(100)
        e[j] = 1.0
        A[:, j] = _apply_xy(oracle, x, y, f, e.reshape(10, 10)).ravel()
    want = np.linalg.solve(A, b)
    assert np.allclose(u, want, rtol=1e-12, atol=0)


def test_implicit_step_conserves_total_u(oracle):
    s = _solver(oracle, 30, coef=2)
    x, y, f = s.x, s.y, s.f
    rro = oracle.cg_init(x, y, H, 2, s.rx, s.ry, f)
    for name in ("u", "p"):
        oracle.halo_field(x, y, H, 1, f[name])
    oracle.copy_u(x, y, H, f)
    total0 = _interior(s, f["u0"]).sum()
    for _ in range(2000):
        pw = oracle.cg_calc_w(x, y, H, f)
        rrn = oracle.cg_calc_ur(x, y, H, rro / pw, f)
        oracle.cg_calc_p(x, y, H, rrn / rro, f)
        for name in ("u", "p"):
            oracle.halo_field(x, y, H, 1, f[name])
        rro = rrn
        if np.sqrt(abs(rrn)) < 1e-15:
            break
    assert np.isclose(_interior(s, f["u"]).sum(), total0, rtol=1e-11)


@pytest.mark.parametrize("depth", [1, 2])
def test_halo_mirrors(oracle, depth):
    x, y = 9, 7
    b = np.arange(x * y, dtype=np.float64)
    oracle.halo_field(x, y, H, depth, b)
    g = b.reshape(y, x)
    for k in range(depth):
        # Interior rows: columns reflect about the face.
        assert np.array_equal(g[H:y - H, H - 1 - k], g[H:y - H, H + k])
        assert np.array_equal(g[H:y - H, x - H + k], g[H:y - H, x - H - 1 - k])
        # All columns (corners included): rows reflect about the face.
        assert np.array_equal(g[H - 1 - k, :], g[H + k, :])
        assert np.array_equal(g[y - H + k, :], g[y - H - 1 - k, :])
    # Cells outside the updated depth are untouched.
    if depth == 1:
        assert g[H, 0] == H * x and g[0, H] == H


def test_field_summary_closed_form(oracle):
    # On 20x20 the tea.in rectangles fall on cell edges: low-density area 16
    # (state 2: 1x1 at energy 25; states 3-5: 15 at energy 0.1), the rest 84
    # at density 100, energy 1e-4.  temp = sum V*rho*u with u = energy*rho.
    s = _solver(oracle, 20)
    vol, mass, ie, temp = oracle.field_summary(s.x, s.y, H, s.f)
    assert np.isclose(vol, 100.0, rtol=1e-14)
    assert np.isclose(mass, 84 * 100.0 + 16 * 0.1, rtol=1e-14)
    assert np.isclose(ie, 84 * 100.0 * 1e-4 + 0.1 * 25.0 + 15 * 0.1 * 0.1, rtol=1e-14)
    assert np.isclose(temp, 84 * 1e4 * 1e-4 + 0.01 * 25.0 + 15 * 0.01 * 0.1,
                      rtol=1e-14)


def test_smvp_single_cell_by_hand(oracle):
    # 3x3 stencil around i with kx=ky=1 everywhere: A a = 5 a_i - sum(neighbours).
    x = 5
    a = np.arange(25, dtype=np.float64) ** 2
    k = np.ones(25)
    i = 12
    want = 5 * a[i] - (a[i + 1] + a[i - 1] + a[i + x] + a[i - x])
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    got = oracle.lib.orc_smvp(a.ctypes.data_as(P), k.ctypes.data_as(P),
                              k.ctypes.data_as(P), i, x)
    assert got == want


@pytest.mark.parametrize("depth", [1, 2])
def test_pack_unpack_strips_are_the_face_slices(oracle, depth):
    """pack_halos.cpp's linear offsets name, for each face, the `depth`
    columns (rows) next to the interior (pack) or in the halo (unpack), row
    by row: the 2-D slices below."""
    x, y = 13, 9
    rng = np.random.default_rng(depth)
    field = rng.standard_normal(x * y)
    g = field.reshape(y, x)
    want = {0: g[:, H:H + depth], 1: g[:, x - H - depth:x - H],
            2: g[H:H + depth, :], 3: g[y - H - depth:y - H, :]}
    for face, sl in want.items():
        buf = np.zeros(sl.size)
        oracle.pack_or_unpack(x, y, H, depth, face, True, field, buf)
        assert np.array_equal(buf, sl.ravel()), face
    halo = {0: (slice(None), slice(H - depth, H)), 1: (slice(None), slice(x - H, x - H + depth)),
            2: (slice(H - depth, H), slice(None)), 3: (slice(y - H, y - H + depth), slice(None))}
    for face, idx in halo.items():
        f2 = field.copy()
        buf = rng.standard_normal(g[idx].size)
        oracle.pack_or_unpack(x, y, H, depth, face, False, f2, buf)
        expect = g.copy()
        expect[idx] = buf.reshape(expect[idx].shape)
        assert np.array_equal(f2.reshape(y, x), expect), face
