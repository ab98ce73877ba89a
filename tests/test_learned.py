"""Host side of the learned preconditioner: the training problems' fields and
coefficients against the oracle's cg_init, the tile decoupling used to batch
them, and the network file."""

import numpy as np
import pytest

from exploringsycl_amd import learned
from tests.oracle_ctypes import new_fields

H = 2


def test_tile_problems_are_seeded_and_in_range():
    a = learned.tile_problem(7)
    b = learned.tile_problem(7)
    for u, v in zip(a[:2], b[:2]):
        assert np.array_equal(u, v)
    dens, ener, dx = a
    assert dens.shape == (256, 256)
    assert dx == 10.0 / learned.RESOLUTIONS[7 % 2]
    rect = dens != 100.0
    assert rect.any()
    lo, hi = learned.DENSITY_RANGE
    assert dens[rect].min() >= 10.0 ** lo and dens[rect].max() <= 10.0 ** hi
    assert (ener[rect] >= 0.1).all() and (ener[rect] <= 25.0).all()
    assert not np.array_equal(learned.tile_problem(8)[0], dens)


def _oracle_init(oracle, dens, dx):
    ny, nx = dens.shape
    x, y = nx + 2 * H, ny + 2 * H
    f = new_fields(x, y, H)
    f["density"].reshape(y, x)[H:-H, H:-H] = dens
    f["energy"].reshape(y, x)[H:-H, H:-H] = 1.0
    for k in ("density", "energy"):
        oracle.halo_field(x, y, H, 2, f[k])
    rx = learned.DT / (dx * dx)
    oracle.cg_init(x, y, H, 2, rx, rx, f)
    return x, y, f


def test_face_coefficients_match_oracle_cg_init(oracle):
    dens, _, dx = learned.tile_problem(3)
    kx, ky = learned.face_coefficients(dens, dx)
    x, y, f = _oracle_init(oracle, dens, dx)
    okx = f["kx"].reshape(y, x)[H:y - H, H:x - H + 1]
    oky = f["ky"].reshape(y, x)[H:y - H + 1, H:x - H]
    assert np.array_equal(kx, okx)
    assert np.array_equal(ky, oky)


def test_zero_faces_decouple_tiles(oracle):
    """A 2 x 2 tiling with zero faces on every tile edge applies, tile by
    tile, the operator of each tile alone with reflective boundaries."""
    rng = np.random.default_rng(5)
    n, g = 16, 2
    dx = 10.0 / 4096
    tiles = [10.0 ** rng.uniform(-2, 2, (n, n)) for _ in range(g * g)]
    X = g * n + 2 * H
    kx = np.zeros((X, X))
    ky = np.zeros((X, X))
    p = np.zeros((X, X))
    want = np.zeros((g * n, g * n))
    for t, dens in enumerate(tiles):
        gy, gx = divmod(t, g)
        fx, fy = learned.face_coefficients(dens, dx)
        fx[:, [0, -1]] = 0.0
        fy[[0, -1], :] = 0.0
        kx[H + gy * n:H + (gy + 1) * n, H + gx * n:H + (gx + 1) * n + 1] = fx
        ky[H + gy * n:H + (gy + 1) * n + 1, H + gx * n:H + (gx + 1) * n] = fy
        v = rng.standard_normal((n, n))
        p[H + gy * n:H + (gy + 1) * n, H + gx * n:H + (gx + 1) * n] = v
        # The same tile alone, reflective halos, coefficients from cg_init.
        x, y, f = _oracle_init(oracle, dens, dx)
        f["p"].reshape(y, x)[H:-H, H:-H] = v
        oracle.halo_field(x, y, H, 1, f["p"])
        oracle.cg_calc_w(x, y, H, f)
        want[gy * n:(gy + 1) * n, gx * n:(gx + 1) * n] = f["w"].reshape(y, x)[H:-H, H:-H]
    f = dict(p=p.ravel().copy(), w=np.zeros(X * X), kx=kx.ravel(), ky=ky.ravel())
    oracle.cg_calc_w(X, X, H, f)
    got = f["w"].reshape(X, X)[H:-H, H:-H]
    assert np.allclose(got, want, rtol=1e-12, atol=0)


def test_network_file_round_trip(tmp_path):
    params = np.random.default_rng(0).standard_normal(38177).astype(np.float32)
    norm = np.array([1.5, 2.0, -0.5, 3.0], dtype=np.float32)
    path = str(tmp_path / "net.bin")
    learned.write_weights(path, params, norm)
    p2, n2, c = learned.read_weights(path)
    assert np.array_equal(p2, params) and np.array_equal(n2, norm) and c == learned.COARSEN
    with open(path, "r+b") as f:
        f.truncate(1000)
    with pytest.raises(ValueError):
        learned.read_weights(path)


def test_initial_params_layout():
    p = learned.initial_params(0)
    assert p.size == 38177 and p.dtype == np.float32
    # Biases are zero: the last parameter is b6, the 32 before W6 are b5.
    assert p[-1] == 0.0
    assert np.all(p[-1 - 288 - 32:-1 - 288] == 0.0)


def test_coarse_faces_are_the_galerkin_operator():
    """coarse_faces: I + the 5-point operator with its face coefficients is
    P^T A P / c^2 for the piecewise-constant P, A the tile's operator with
    reflective edges (zero edge faces)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(2)
    n, c = 32, 8
    dens = 10.0 ** rng.uniform(-2, 2, (n, n))
    kx, ky = learned.face_coefficients(dens, 10.0 / 4096)
    kx = kx.astype(np.float32).astype(np.float64)
    ky = ky.astype(np.float32).astype(np.float64)
    kx[:, [0, -1]] = 0.0
    ky[[0, -1], :] = 0.0

    def five_point(kx, ky):
        m = ky.shape[1]
        N = kx.shape[0] * m
        A = sp.lil_matrix((N, N))
        for j in range(kx.shape[0]):
            for k in range(m):
                i = j * m + k
                A[i, i] = 1 + kx[j, k] + kx[j, k + 1] + ky[j, k] + ky[j + 1, k]
                if k > 0:
                    A[i, i - 1] = -kx[j, k]
                if k < m - 1:
                    A[i, i + 1] = -kx[j, k + 1]
                if j > 0:
                    A[i, i - m] = -ky[j, k]
                if j < kx.shape[0] - 1:
                    A[i, i + m] = -ky[j + 1, k]
        return A.tocsr()

    A = five_point(kx, ky)
    rows = np.arange(n * n)
    cols = (rows // n // c) * (n // c) + (rows % n) // c
    P = sp.csr_matrix((np.ones(n * n), (rows, cols)), shape=(n * n, (n // c) ** 2))
    want = (P.T @ A @ P).toarray() / (c * c)
    kxc, kyc = learned.coarse_faces(kx, ky, c)
    got = five_point(kxc, kyc).toarray()
    assert np.allclose(got, want, rtol=1e-13, atol=1e-12)


def _coarse_chunk(n=12, seed=3):
    rng = np.random.default_rng(seed)
    X = n + 4
    kx = np.zeros((X, X))
    ky = np.zeros((X, X))
    kx[2:2 + n, 3:2 + n] = 10.0 ** rng.uniform(-1, 3, (n, n - 1))
    ky[3:2 + n, 2:2 + n] = 10.0 ** rng.uniform(-1, 3, (n - 1, n))
    return kx, ky


def _dense(kx, ky, n):
    A = np.zeros((n * n, n * n))
    for j in range(n):
        for k in range(n):
            i = j * n + k
            J, K = j + 2, k + 2
            A[i, i] = 1 + kx[J, K] + kx[J, K + 1] + ky[J, K] + ky[J + 1, K]
            if k > 0:
                A[i, i - 1] = -kx[J, K]
            if k < n - 1:
                A[i, i + 1] = -kx[J, K + 1]
            if j > 0:
                A[i, i - n] = -ky[J, K]
            if j < n - 1:
                A[i, i + n] = -ky[J + 1, K]
    return A


def test_jacobi_sweeps_pinned():
    """jacobi_sweeps: one sweep from zero is rhs / diag; many sweeps reach
    the dense solve of (I + L) e = rhs; 6 sweeps are a symmetric positive
    definite linear map whose reach is exactly 6 cells."""
    n = 12
    kx, ky = _coarse_chunk(n)
    A = _dense(kx, ky, n)
    rng = np.random.default_rng(4)
    rhs = np.zeros((n + 4, n + 4))
    rhs[2:-2, 2:-2] = rng.standard_normal((n, n))
    one = learned.jacobi_sweeps(rhs, kx, ky, 1)[2:-2, 2:-2].ravel()
    assert np.allclose(one, rhs[2:-2, 2:-2].ravel() / np.diag(A), rtol=1e-14, atol=0)
    many = learned.jacobi_sweeps(rhs, kx, ky, 20000)[2:-2, 2:-2].ravel()
    assert np.allclose(many, np.linalg.solve(A, rhs[2:-2, 2:-2].ravel()), rtol=1e-8, atol=1e-12)
    # The 6-sweep map as a matrix, p(D^-1 T) D^-1 = D^-1/2 p(S) D^-1/2 with
    # S = D^-1/2 T D^-1/2: symmetric, positive definite, zero beyond 6 cells.
    M = np.zeros((n * n, n * n))
    for q in range(n * n):
        e = np.zeros((n + 4, n + 4))
        e[2 + q // n, 2 + q % n] = 1.0
        M[:, q] = learned.jacobi_sweeps(e, kx, ky, 6)[2:-2, 2:-2].ravel()
    assert np.allclose(M, M.T, rtol=1e-10, atol=1e-14)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    jj, kk = np.divmod(np.arange(n * n), n)
    far = (np.abs(jj[:, None] - jj[None, :]) + np.abs(kk[:, None] - kk[None, :])) > 6
    assert np.all(M[far] == 0.0)


def test_local_scale_is_the_window_rms():
    """local_scale: the RMS over the 13 x 13 window with zero padding, as a
    brute-force loop computes it."""
    v = np.random.default_rng(6).standard_normal((20, 17))
    got = learned.local_scale(v)
    w = learned.WINDOW
    want = np.zeros_like(v)
    for j in range(v.shape[0]):
        for k in range(v.shape[1]):
            blk = v[max(0, j - w):j + w + 1, max(0, k - w):k + w + 1]
            want[j, k] = np.sqrt((blk * blk).sum() / (2 * w + 1) ** 2)
    assert np.allclose(got, want, rtol=1e-12, atol=0)
