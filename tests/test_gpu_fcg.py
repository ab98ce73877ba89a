"""Flexible CG on the GPU: the learned preconditioner's application against
a numpy restatement with the network oracle, the training-pair generator,
convergence of the learned solver, and what the trained network gains."""

import os
import shutil

import numpy as np
import pytest

from exploringsycl_amd import learned
from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import run_deck
from oracle import nn_oracle as O

pytestmark = pytest.mark.gpu
H = 2


def _fcg_chunk(n, seed, params, norm):
    from exploringsycl_amd.chunk import DeviceChunk
    rng = np.random.default_rng(seed)
    X = n + 2 * H
    c = DeviceChunk(X, X, H)
    dens = np.full((X, X), 100.0)
    dens[H:-H, H:-H] = 10.0 ** rng.uniform(-2, 2, (n, n))
    dens = np.pad(dens[H:-H, H:-H], H, mode="symmetric")
    ener = rng.uniform(0.1, 25.0, (X, X))
    c.write("density", dens)
    c.write("energy", ener)
    c.fcg_attach(params, norm, learned.COARSEN)
    dx = 10.0 / 4096
    rx = learned.DT / (dx * dx)
    rr, bb = c.fcg_init(2, rx, rx)
    return c, rr, bb


def _bf(v):
    return O.bf16(np.asarray(v, dtype=np.float32)).astype(np.float64)


@pytest.mark.parametrize("n", [32 * learned.COARSEN, 64 * learned.COARSEN])
def test_learned_preconditioner_matches_numpy(n):
    """z = r/diag + s P(N(R r/s, f)) and z.w against restrict + the network
    oracle + prolong in numpy on the device's own FP32 r, kx, ky, w."""
    params = learned.initial_params(3)
    # Non-zero biases so every term is exercised.
    rng = np.random.default_rng(4)
    k = 0
    for o, i in O.SHAPES:
        k += o * i * 9
        params[k:k + o] = rng.uniform(-0.05, 0.05, o)
        k += o
    norm = np.array([3.0, 2.0, 3.5, 2.5], dtype=np.float32)
    c, _, _ = _fcg_chunk(n, n, params, norm)
    try:
        # One direction, p = Jacobi(r) and w = A p, so that w is not zero.
        c.fcg_precondition(False)
        c.fcg_calc_w(0.0, False, 0.0)
        zw = c.fcg_precondition(True)
        z = c.read_z().astype(np.float64)
        r = c.read_mp("r").astype(np.float64)
        kx = c.read_mp("kx").astype(np.float64)
        ky = c.read_mp("ky").astype(np.float64)
        w = c.read_mp("w").astype(np.float64)
    finally:
        c.close()
    cc = learned.COARSEN
    ri = r[H:-H, H:-H]
    rbar = ri.reshape(n // cc, cc, n // cc, cc).sum(axis=(1, 3)) / (cc * cc)
    rbar = rbar.astype(np.float32).astype(np.float64)
    s = np.float32(np.sqrt((rbar * rbar).sum() / rbar.size))
    kxc, kyc = learned.coarse_faces(kx[H:H + n, H:H + n + 1], ky[H:H + n + 1, H:H + n], cc)
    g0, g1 = learned.coarse_features(kxc, kyc)
    f0 = (g0 - norm[0]) / norm[1]
    f1 = (g1 - norm[2]) / norm[3]
    x = np.stack([np.float32(rbar) * (np.float32(1) / s), f0, f1], axis=-1)[None]
    out, _ = O.forward(params, x.astype(np.float32), bf16_mode=True)
    out = out[0].astype(np.float64)
    e = H + n  # one past the last interior index
    d = (1.0 + (kx[H:e, H + 1:e + 1] + kx[H:e, H:e])) + (ky[H + 1:e + 1, H:e] + ky[H:e, H:e])
    want = ri / d + float(s) * np.kron(out, np.ones((cc, cc)))
    got = z[H:-H, H:-H]
    corr = float(s) * np.abs(out).max()
    assert np.abs(got - want).max() <= 1e-2 * corr + 1e-6 * np.abs(want).max()
    want_zw = (got * w[H:-H, H:-H]).sum()
    assert abs(zw - want_zw) <= 1e-9 * np.abs(got * w[H:-H, H:-H]).sum()


def test_training_pairs_satisfy_the_operator(oracle):
    """Pairs from the tiled FP64 CG on the coarse operators: T e_k = r_k on
    the tile alone, T applied by the oracle's stencil with the tile's coarse
    face coefficients (zero on its edges)."""
    data = learned.generate(0, 3, snapshots=2, rng_seed=1, keep_coarse=True)
    n = data["inputs"].shape[0]
    cw = learned.TILE // learned.COARSEN
    assert n >= 3 and data["targets"].shape == (n, cw, cw)
    assert np.isfinite(data["inputs"]).all() and np.isfinite(data["targets"]).all()
    r, e = data["coarse_r"], data["coarse_e"]
    X = cw + 2 * H
    kx = np.zeros((X, X))
    ky = np.zeros((X, X))
    kx[H:H + cw, H:H + cw + 1] = data["coarse_kx"]
    ky[H:H + cw + 1, H:H + cw] = data["coarse_ky"]
    p = np.zeros((X, X))
    p[H:-H, H:-H] = e
    f = dict(p=p.ravel().copy(), w=np.zeros(X * X), kx=kx.ravel(), ky=ky.ravel())
    oracle.halo_field(X, X, H, 1, f["p"])
    oracle.cg_calc_w(X, X, H, f)
    Te = f["w"].reshape(X, X)[H:-H, H:-H]
    assert np.linalg.norm(Te - r) <= 1e-6 * np.linalg.norm(r)


def _fcg_deck(n, seed, extra):
    return Deck.with_random_rectangles(
        n, n, 64, problem_seed(5, seed), density="log10", density_range=(-2.0, 2.0),
        end_step=1, coefficient="inverse", extent=10.0 * n / 4096,
        extra=["rel_eps=1e-10"] + extra)


def _run_fcg(tmp_path, n, seed, weights, name):
    wd = tmp_path / name
    wd.mkdir()
    extra = ["use_jacobi_pcg"]
    if weights is not None:
        shutil.copy(weights, wd / "net.bin")
        extra = ["use_fcg", "fcg_weights=net.bin"]
    rec, _ = run_deck(_fcg_deck(n, seed, extra), workdir=str(wd))
    return rec


def test_fcg_with_untrained_network_converges(tmp_path):
    """A random network is a poor preconditioner; flexible CG (with its
    Jacobi fallback) must still reach 1e-10 and the FP64 CG answer."""
    path = str(tmp_path / "random.bin")
    learned.write_weights(path, learned.initial_params(11),
                          np.array([3.0, 2.0, 3.0, 2.0], dtype=np.float32))
    rec = _run_fcg(tmp_path, 256, 1, path, "fcg")
    assert rec["solver"] == "FCG"
    s = rec["steps"][0]
    assert s["rel_residual"] < 1e-10
    ref, _ = run_deck(_fcg_deck(256, 1, ["rel_eps=1e-13"]), workdir=str(tmp_path / "cg"))
    a, b = rec["summaries"][-1], ref["summaries"][-1]
    assert abs(a["temp"] / b["temp"] - 1) < 1e-8


@pytest.mark.skipif(not os.path.exists(learned.WEIGHTS_PATH),
                    reason="no trained network file")
@pytest.mark.parametrize("n", [512, 1024])
def test_trained_network_gains_over_jacobi(tmp_path, n):
    """The trained network against the same flexible CG without it (Jacobi
    only) on unseen fields: fewer iterations to the same 1e-10, and no
    fallback."""
    learned_rec = _run_fcg(tmp_path, n, 900 + n, learned.WEIGHTS_PATH, "fcg")
    jacobi_rec = _run_fcg(tmp_path, n, 900 + n, None, "jpcg")
    a, b = learned_rec["steps"][0], jacobi_rec["steps"][0]
    print(f"{n}^2: learned FCG {a['iterations']} iterations, Jacobi-PCG "
          f"{b['iterations']}")
    assert a["rel_residual"] < 1e-10 and b["rel_residual"] < 1e-10
    assert a["fallback_at"] == -1
    assert a["iterations"] < b["iterations"]
