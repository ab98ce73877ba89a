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


def _fcg_chunk(n, seed, params, norm, levels=2):
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
    c.fcg_attach(params, norm, learned.COARSEN, levels)
    dx = 10.0 / 4096
    rx = learned.DT / (dx * dx)
    rr, bb = c.fcg_init(2, rx, rx)
    return c, rr, bb


def _bf(v):
    return O.bf16(np.asarray(v, dtype=np.float32)).astype(np.float64)


def _restated_levels(params, norm, r, kx, ky, n, levels):
    """The learned preconditioner's coarse terms restated in numpy on the
    device's own FP32 r, kx, ky: per level l (coarsening c = COARSEN * 4^l
    while the coarse width stays a multiple of 32) the block means, local
    scale, network oracle, unscaling and energy-optimal factor g; returns
    [(c, g, y)]."""
    out = []
    prev = None
    for l in range(levels):
        cc = learned.COARSEN << (2 * l)
        if n % cc or (n // cc) % 32:
            break
        m = n // cc
        if prev is None:
            ri = r[H:-H, H:-H]
            rbar = ri.reshape(m, cc, m, cc).sum(axis=(1, 3)) / (cc * cc)
        else:
            rbar = prev.reshape(m, 4, m, 4).sum(axis=(1, 3)) / 16.0
        rbar = rbar.astype(np.float32).astype(np.float64)
        prev = rbar
        kxc, kyc = learned.coarse_faces(kx[H:H + n, H:H + n + 1], ky[H:H + n + 1, H:H + n], cc)
        dis = learned.coarse_scaling(kxc, kyc).astype(np.float32)
        rs = (np.float32(rbar) * dis).astype(np.float64)
        sig = np.maximum(learned.local_scale(rs), learned.FLOOR * np.sqrt(np.mean(rs * rs)))
        sig = sig.astype(np.float32)
        g0, g1 = learned.coarse_features(kxc, kyc)
        f0 = (g0 - norm[0]) / norm[1]
        f1 = (g1 - norm[2]) / norm[3]
        x = np.stack([(np.float32(rbar) * dis) / sig, f0, f1], axis=-1)[None]
        o, _ = O.forward(params, x.astype(np.float32), bf16_mode=True)
        y = (o[0] * (sig * dis)).astype(np.float64)
        kl = kxc[:, :-1].astype(np.float32).astype(np.float64)
        kr = kxc[:, 1:].astype(np.float32).astype(np.float64)
        kb = kyc[:-1, :].astype(np.float32).astype(np.float64)
        kt = kyc[1:, :].astype(np.float32).astype(np.float64)
        def T(v, kl=kl, kr=kr, kb=kb, kt=kt):
            vp = np.pad(v, 1)
            return ((1.0 + kl + kr + kb + kt) * v - kl * vp[1:-1, :-2] - kr * vp[1:-1, 2:]
                    - kb * vp[:-2, 1:-1] - kt * vp[2:, 1:-1])
        out.append(dict(c=cc, y=y, a=cc * cc * (y * rbar).sum(), b=cc * cc * (y * T(y)).sum(),
                        T=T))
    if len(out) == 2:
        # Joint factors: the 2 x 2 Galerkin system of the two directions.
        l1, l2 = out
        qy2 = np.kron(l2["y"], np.ones((4, 4)))
        b12 = l1["c"] ** 2 * (l1["y"] * l1["T"](qy2)).sum()
        det = l1["b"] * l2["b"] - b12 * b12
        g = np.linalg.solve([[l1["b"], b12], [b12, l2["b"]]], [l1["a"], l2["a"]])
        assert det > 1e-12 * l1["b"] * l2["b"]
        return [(l1["c"], g[0], l1["y"]), (l2["c"], g[1], l2["y"])]
    return [(t["c"], t["a"] / t["b"], t["y"]) for t in out]


@pytest.mark.parametrize("n,levels", [(32 * learned.COARSEN, 2), (64 * learned.COARSEN, 1),
                                      (128 * learned.COARSEN, 2)])
def test_learned_preconditioner_matches_numpy(n, levels):
    """z = r/diag + sum over levels of g P y and z.w against the numpy
    restatement with the network oracle, on the device's own FP32 r, kx,
    ky, w (256^2 and 512^2 have one level, 1024^2 two)."""
    params = learned.initial_params(3)
    # Non-zero biases so every term is exercised.
    rng = np.random.default_rng(4)
    k = 0
    for o, i in O.SHAPES:
        k += o * i * 9
        params[k:k + o] = rng.uniform(-0.05, 0.05, o)
        k += o
    norm = np.array([0.2, 0.1, 0.2, 0.1], dtype=np.float32)
    c, _, _ = _fcg_chunk(n, n, params, norm, levels)
    try:
        used = c.fcg_levels()
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
    terms = _restated_levels(params, norm, r, kx, ky, n, levels)
    assert used == len(terms)
    ri = r[H:-H, H:-H]
    e = H + n  # one past the last interior index
    d = (1.0 + (kx[H:e, H + 1:e + 1] + kx[H:e, H:e])) + (ky[H + 1:e + 1, H:e] + ky[H:e, H:e])
    want = ri / d
    corr = 0.0
    for cc, gamma, y in terms:
        want = want + float(np.float32(gamma)) * np.kron(y, np.ones((cc, cc)))
        corr += abs(gamma) * np.abs(y).max()
    got = z[H:-H, H:-H]
    assert np.abs(got - want).max() <= 2e-2 * corr + 1e-6 * np.abs(want).max()
    want_zw = (got * w[H:-H, H:-H]).sum()
    assert abs(zw - want_zw) <= 1e-9 * np.abs(got * w[H:-H, H:-H]).sum()


def test_training_pairs():
    """Training pairs from the generator on the GPU: each target is the
    Jacobi sweeps of its input field on the problems' coarse operators
    (zero faces between problems), and the first pair's input channels,
    features and target are that field, those coefficients and those
    sweeps in the scaled variables over the input's local scale."""
    data = learned.generate(0, 3, snapshots=2, rng_seed=1, keep_coarse=True)
    n = data["inputs"].shape[0]
    cw = learned.TILE // learned.COARSEN
    assert n >= 3 and data["targets"].shape == (n, cw, cw)
    assert np.isfinite(data["inputs"]).all() and np.isfinite(data["targets"]).all()
    field, ek = data["coarse_field"], data["coarse_e_field"]
    want = learned.jacobi_sweeps(field, data["chunk_kx"], data["chunk_ky"])
    assert np.array_equal(ek, want)
    # The first pair: tile (0, 0) of the first field.
    ys = xs = slice(H, H + cw)
    kxc = data["chunk_kx"][ys, H:H + cw + 1]
    kyc = data["chunk_ky"][H:H + cw + 1, xs]
    dis = learned.coarse_scaling(kxc, kyc)
    rs = field[ys, xs] * dis
    sig = np.maximum(learned.local_scale(rs), learned.FLOOR * np.sqrt(np.mean(rs * rs)))
    assert np.allclose(data["inputs"][0, ..., 0], rs / sig, rtol=1e-6)
    assert np.allclose(data["targets"][0], ek[ys, xs] / dis / sig, rtol=1e-6)
    f0, f1 = learned.coarse_features(kxc, kyc)
    assert np.allclose(data["inputs"][0, ..., 1], f0, rtol=1e-6)
    assert np.allclose(data["inputs"][0, ..., 2], f1, rtol=1e-6)


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
                          np.array([0.2, 0.1, 0.2, 0.1], dtype=np.float32))
    rec = _run_fcg(tmp_path, 256, 1, path, "fcg")
    assert rec["solver"] == "FCG"
    s = rec["steps"][0]
    assert s["rel_residual"] < 1e-10
    ref, _ = run_deck(_fcg_deck(256, 1, ["rel_eps=1e-13"]), workdir=str(tmp_path / "cg"))
    a, b = rec["summaries"][-1], ref["summaries"][-1]
    assert abs(a["temp"] / b["temp"] - 1) < 1e-8


@pytest.mark.skipif(not os.path.exists(learned.WEIGHTS_PATH),
                    reason="no trained network file")
@pytest.mark.parametrize("n,seed", [(512, 3000), (4096, 3000)])
def test_trained_network_gains_over_jacobi(tmp_path, n, seed):
    """What the trained network gains: flexible CG with it against the same
    flexible CG without it (Jacobi-PCG) on an unseen field of the
    64-rectangle workload (log10 density in [-2, 2]), both to 1e-10 relative
    residual; the learned solve must need at most 2/3 of the iterations,
    without falling back, and agree with the Jacobi solve's temperature."""
    learned_rec = _run_fcg(tmp_path, n, seed, learned.WEIGHTS_PATH, "fcg")
    jacobi_rec = _run_fcg(tmp_path, n, seed, None, "jpcg")
    a, b = learned_rec["steps"][0], jacobi_rec["steps"][0]
    print(f"{n}^2: learned FCG {a['iterations']} iterations, Jacobi-PCG "
          f"{b['iterations']}")
    assert a["rel_residual"] < 1e-10 and b["rel_residual"] < 1e-10
    assert a["fallback_at"] == -1
    assert a["iterations"] <= (2 * b["iterations"]) // 3
    ta, tb = learned_rec["summaries"][-1]["temp"], jacobi_rec["summaries"][-1]["temp"]
    assert abs(ta / tb - 1) < 1e-8


# ---- the two-level preconditioner with Jacobi sweeps as the coarse solve ----

def _sweep_deck(n, seed, extra, count=24, halo=None):
    more = [f"halo_depth={halo}"] if halo else []
    return Deck.with_random_rectangles(
        n, n, count, problem_seed(5, seed), density="log10", density_range=(-2.0, 2.0),
        end_step=1, coefficient="inverse", extent=10.0 * n / 4096,
        extra=["rel_eps=1e-10"] + extra + more)


def _same_run(tmp_path, deck):
    """Runs the deck on the GPU and on the oracle host; asserts the same
    iteration count, fallback, summaries and final field, bit for bit."""
    from tests.oracle_ctypes import ORACLE_TEALEAF
    gpu, _ = run_deck(deck, workdir=str(tmp_path / "gpu"), extra_args=["--field-out", "u.bin"])
    cpu, _ = run_deck(deck, binary=ORACLE_TEALEAF, workdir=str(tmp_path / "cpu"),
                      extra_args=["--field-out", "u.bin"])
    a, b = gpu["steps"][0], cpu["steps"][0]
    for k in ("iterations", "executed", "refreshes", "fallback_at", "rel_residual"):
        assert a[k] == b[k], (k, a[k], b[k])
    assert gpu["summaries"] == cpu["summaries"]
    ug = np.fromfile(tmp_path / "gpu" / "u.bin")
    uc = np.fromfile(tmp_path / "cpu" / "u.bin")
    assert np.array_equal(ug, uc)
    return gpu


@pytest.mark.parametrize("n,sweeps,levels,halo", [
    (64, 6, 2, None), (96, 6, 2, None), (128, 1, 2, None), (64, 3, 1, None),
    (200, 6, 2, None), (256, 8, 2, None), (64, 6, 2, 3), (512, 6, 2, None)])
def test_sweeps_fcg_matches_oracle_bitwise(tmp_path, n, sweeps, levels, halo):
    """Flexible CG with the non-learned two-level preconditioner (restriction
    of both levels, Jacobi sweeps on the coarse operators, the joint factors,
    the prolongation) against the oracle's restatement, the whole solve:
    every sum is in a fixed order, so iteration counts, summaries and the
    final field agree bit for bit (odd x with halo depth 3; 200^2 has one
    level, its 25 x 25 level 0 not a multiple of the 16-cell tile)."""
    deck = _sweep_deck(n, n + sweeps, ["use_fcg", f"fcg_coarse_sweeps={sweeps}",
                                        f"fcg_levels={levels}"], halo=halo)
    rec = _same_run(tmp_path, deck)
    assert rec["coarse_solve"] == f"jacobi sweeps {sweeps}"
    assert rec["steps"][0]["rel_residual"] < 1e-10


def test_sweeps_fcg_fallback_matches_oracle_bitwise(tmp_path):
    """A stall rule no method meets: the fallback fires at the end of the
    first window on the GPU as on the oracle, and both continue as the same
    Jacobi-PCG to the same field."""
    deck = _sweep_deck(96, 11, ["use_fcg", "fcg_coarse_sweeps=6", "stall_window=20",
                                "stall_factor=1000"])
    rec = _same_run(tmp_path, deck)
    assert rec["steps"][0]["fallback_at"] == 20
    assert rec["steps"][0]["fallback_beta"] == 0.0


def test_learned_fcg_falls_back_and_converges(tmp_path):
    """The learned preconditioner under a stall rule no method meets (and
    under the specification's literal 2x per 20 iterations): the fallback
    fires, the recursion restarts with beta = 0, and the solve reaches 1e-10
    at the Jacobi-PCG temperature."""
    if not os.path.exists(learned.WEIGHTS_PATH):
        pytest.skip("no trained network file")
    temps = {}
    for name, rule in (("strict", ["stall_window=20", "stall_factor=1000"]),
                       ("literal", ["stall_window=20", "stall_factor=2"])):
        wd = tmp_path / name
        wd.mkdir()
        shutil.copy(learned.WEIGHTS_PATH, wd / "net.bin")
        rec, _ = run_deck(_sweep_deck(512, 5, ["use_fcg", "fcg_weights=net.bin"] + rule,
                                      count=64), workdir=str(wd))
        st = rec["steps"][0]
        assert st["fallback_at"] > 0 and st["fallback_beta"] == 0.0
        if name == "strict":
            assert st["fallback_at"] == 20
        assert st["rel_residual"] < 1e-10
        temps[name] = rec["summaries"][-1]["temp"]
    jp, _ = run_deck(_sweep_deck(512, 5, ["use_jacobi_pcg"], count=64),
                     workdir=str(tmp_path / "jp"))
    for t in temps.values():
        assert abs(t / jp["summaries"][-1]["temp"] - 1) < 1e-8


@pytest.mark.skipif(not os.path.exists(learned.WEIGHTS_PATH), reason="no trained network file")
@pytest.mark.parametrize("n", [512, 1024, 2048])
def test_fused_network_level_is_bitwise_the_per_layer_pipeline(tmp_path, n):
    """The learned coarse solve in one launch per level (net_level_kernel,
    TEA_FCG_NET=fused) against the per-layer pipeline (box_rows, box_cols,
    the network's own layer kernels, net_tile_kernel): every output pixel
    takes the same MFMA sequence, so the whole solve -- iteration count,
    summaries and final field -- is the same bit for bit (512^2 has one
    coarse level, 1024^2 and 2048^2 two)."""
    runs = {}
    for mode in ("fused", "layers"):
        wd = tmp_path / mode
        wd.mkdir()
        shutil.copy(learned.WEIGHTS_PATH, wd / "net.bin")
        env = {"TEA_FCG_NET": mode}
        rec, _ = run_deck(_fcg_deck(n, 7, ["use_fcg", "fcg_weights=net.bin"]),
                          workdir=str(wd), env=env, extra_args=["--field-out", "u.bin"])
        runs[mode] = (rec, np.fromfile(wd / "u.bin"))
    (a, ua), (b, ub) = runs["fused"], runs["layers"]
    assert a["steps"][0]["iterations"] == b["steps"][0]["iterations"]
    assert a["steps"][0]["rel_residual"] == b["steps"][0]["rel_residual"]
    assert a["steps"][0]["rel_residual"] < 1e-10
    assert a["summaries"] == b["summaries"]
    assert np.array_equal(ua, ub)
