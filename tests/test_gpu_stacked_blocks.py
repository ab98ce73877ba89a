"""Stacked blocks in the stencil passes: with TEA_HIP_STACK=K a workgroup of
the four-cells-per-lane mixed-precision w-step (mp_w4_kernel, also FCG's),
and of the pair-per-thread FP64 CG pass
(fused_w1_kernel) and Jacobi sweep (jacobi_pair_kernel, checked against the
oracle in test_gpu_solvers.py) takes K blocks one row apart
(block c + S t, S = x / 512), so their rows above and below come from the
CU's L1.  Every block's cells, lanes and wave sums are unchanged, so the
iterations must give the same scalars and fields bit for bit with K = 1, 2
(the default) and 4, on row lengths whose offset between block rows is 4 and
88 cells and on grids whose block count leaves a partial last workgroup."""

import os

import numpy as np
import pytest

from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import run_deck

pytestmark = pytest.mark.gpu

H = 2


def _chunk(x, y, k):
    from exploringsycl_amd.chunk import DeviceChunk
    old = os.environ.get("TEA_HIP_STACK")
    os.environ["TEA_HIP_STACK"] = str(k)
    try:
        return DeviceChunk(x, y, H)
    finally:
        if old is None:
            del os.environ["TEA_HIP_STACK"]
        else:
            os.environ["TEA_HIP_STACK"] = old


def _same(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64),
                          np.asarray(b, dtype=np.float64).view(np.uint64))


def _fields(c, x, y):
    rng = np.random.default_rng(x + 7 * y)
    c.write("density", 10.0 ** rng.uniform(-2, 2, x * y))
    c.write("energy", rng.uniform(0.1, 25.0, x * y))


def _mpcg(x, y, k, iters=14):
    """The mixed-precision driver's sequence with a refresh after iteration
    5: the w-step in all three u-update forms."""
    with _chunk(x, y, k) as c:
        _fields(c, x, y)
        rr, _ = c.mpcg_init(2, 0.8, 1.3)
        s = [rr]
        alpha = beta = 0.0
        pu = pp = False
        for it in range(iters):
            pw, pr = c.mpcg_calc_w(alpha, pu, beta, pp)
            alpha = pr / pw
            rrn, rw = c.mpcg_calc_r(alpha)
            pu = True
            s += [pw, pr, rrn, rw]
            if it == 5:
                a, b, rw = c.mpcg_refresh(alpha, True)
                s += [a, b, rw]
                pu = False
            beta = -rw / pw
            pp = True
        if pu:
            s += list(c.mpcg_refresh(alpha, True))
        return [np.array(s), c.read("u"), c.read_mp("w").astype(np.float64)]


def _fcg(x, y, k, iters=10):
    with _chunk(x, y, k) as c:
        _fields(c, x, y)
        c.fcg_attach()
        rr, _ = c.fcg_init(2, 0.8, 1.3)
        zw = c.fcg_precondition(False)
        s = [rr, zw]
        alpha = beta = 0.0
        pu = False
        for _ in range(iters):
            pw, pr = c.fcg_calc_w(alpha, pu, beta)
            alpha = pr / pw
            rrn, zw = c.fcg_calc_r_precondition(alpha, False)
            pu = True
            s += [pw, pr, rrn, zw]
            beta = -zw / pw
        s += list(c.mpcg_refresh(alpha, True))
        return [np.array(s), c.read("u")]


def _cg(x, y, k, iters=12):
    """The reference's CG call sequence through the fused pass (with the
    run-ahead, which launches the same kernel)."""
    with _chunk(x, y, k) as c:
        _fields(c, x, y)
        rro = c.cg_init(2, 0.8, 1.3)
        c.local_halos(1, ["u", "p"])
        s = [rro]
        for _ in range(iters):
            pw = c.cg_calc_w()
            alpha = rro / pw
            rrn = c.cg_calc_ur(alpha)
            beta = rrn / rro
            c.cg_calc_p(beta)
            c.local_halos(1, ["u", "p"])
            s += [pw, rrn]
            rro = rrn
        s.append(c.cg_calc_w())
        return [np.array(s)] + [c.read(n) for n in ("u", "p", "r", "w")]


@pytest.mark.parametrize("cells", [(512, 61), (596, 70), (2048, 19)])
@pytest.mark.parametrize("k", [2, 4])
def test_stacked_mixed_precision_w_step(cells, k):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    for a, b in zip(_mpcg(x, y, 1), _mpcg(x, y, k)):
        assert _same(a, b)


def test_stacked_fcg_w_step():
    x, y = 600 + 2 * H, 90 + 2 * H
    for a, b in zip(_fcg(x, y, 1), _fcg(x, y, 2)):
        assert _same(a, b)


def _fcg_two_level(x, y, k, iters=8):
    """FCG with the two-level preconditioner (6 Jacobi sweeps on both coarse
    levels) between the w-steps."""
    with _chunk(x, y, k) as c:
        _fields(c, x, y)
        from exploringsycl_amd.chunk import check
        check(c.lib.tea_hip_fcg_attach_sweeps(c.handle, 6, 8, 2), "fcg_attach_sweeps")
        assert c.fcg_levels() == 2
        rr, _ = c.fcg_init(2, 0.8, 1.3)
        zw = c.fcg_precondition(True)
        s = [rr, zw]
        alpha = beta = 0.0
        pu = False
        for _ in range(iters):
            pw, pr = c.fcg_calc_w(alpha, pu, beta)
            alpha = pr / pw
            rrn, zw = c.fcg_calc_r_precondition(alpha, True)
            pu = True
            s += [pw, pr, rrn, zw]
            beta = -zw / pw
        s += list(c.mpcg_refresh(alpha, True))
        return [np.array(s), c.read("u"), c.read_z().astype(np.float64)]


def test_stacked_two_level_fcg():
    x, y = 1024 + 2 * H, 64 + 2 * H
    for a, b in zip(_fcg_two_level(x, y, 1), _fcg_two_level(x, y, 2)):
        assert _same(a, b)


@pytest.mark.parametrize("cells", [(510, 61), (1022, 40)])
@pytest.mark.parametrize("k", [2, 4])
def test_stacked_fp64_cg_pass(cells, k):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    for a, b in zip(_cg(x, y, 1), _cg(x, y, k)):
        assert _same(a, b)


def test_stacked_whole_solves_match(tmp_path):
    """Two timesteps of a 1024^2 rectangle deck through the host program,
    plain CG and mixed-precision CG, with one and two blocks per workgroup:
    the same iterations, summaries and final field, bit for bit."""
    for extra in ([], ["use_mpcg"]):
        deck = Deck.with_random_rectangles(
            1024, 1024, 24, problem_seed(5, 11), density="log10", density_range=(-2.0, 2.0),
            end_step=2, coefficient="inverse", extent=2.5, extra=["rel_eps=1e-10"] + extra)
        runs = {}
        for k in ("1", "2"):
            wd = tmp_path / f"{len(extra)}_{k}"
            wd.mkdir()
            rec, _ = run_deck(deck, workdir=str(wd), env={"TEA_HIP_STACK": k},
                              extra_args=["--field-out", "u.bin"])
            runs[k] = (rec, np.fromfile(wd / "u.bin"))
        (a, ua), (b, ub) = runs["1"], runs["2"]
        assert [s["iterations"] for s in a["steps"]] == [s["iterations"] for s in b["steps"]]
        assert a["summaries"] == b["summaries"]
        assert _same(ua, ub)
