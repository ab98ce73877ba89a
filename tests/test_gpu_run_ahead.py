"""The run-ahead of the fused CG iteration: after cg_calc_w and cg_calc_ur
the library queues the next pass with alpha or beta formed on the device
(the finisher's last block runs the host's final tree passes and divides),
into spare buffers, and the host's next call takes it over when
its scalar is bitwise the device's.  These tests run the CG driver's call
sequence (TeaLeaf/drivers/cg_driver.c:69-124) with the run-ahead on and off
(TEA_HIP_RUN_AHEAD=0) and require the same scalars and fields bit for bit,
also when the host strays from the sequence (another alpha, reads and
reductions between the calls), which must drop the run-ahead."""

import json
import os

import numpy as np
import pytest

from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import run_deck

pytestmark = pytest.mark.gpu

H = 2


def _chunk(x, y, run_ahead):
    from exploringsycl_amd.chunk import DeviceChunk
    old = os.environ.get("TEA_HIP_RUN_AHEAD")
    os.environ["TEA_HIP_RUN_AHEAD"] = "1" if run_ahead else "0"
    try:
        return DeviceChunk(x, y, H)
    finally:
        if old is None:
            del os.environ["TEA_HIP_RUN_AHEAD"]
        else:
            os.environ["TEA_HIP_RUN_AHEAD"] = old


def _fields(x, y, seed):
    rng = np.random.default_rng(seed)
    density = rng.uniform(0.1, 100.0, x * y)
    energy = rng.uniform(0.1, 25.0, x * y)
    return density, energy


def _cg(x, y, seed, run_ahead, iters, stray=None):
    """cg_init, then `iters` iterations of the driver's sequence; `stray(c,
    it, alpha)` may act between the calls and return the alpha to pass.
    Returns the scalars, the final fields and the run-ahead finisher launches."""
    density, energy = _fields(x, y, seed)
    with _chunk(x, y, run_ahead) as c:
        c.timing(True)
        c.write("density", density)
        c.write("energy", energy)
        rro = c.cg_init(2, 0.8, 1.3)
        c.local_halos(1, ["u", "p"])
        scalars = [rro]
        for it in range(iters):
            pw = c.cg_calc_w()
            alpha = rro / pw
            if stray is not None:
                alpha = stray(c, it, alpha)
            rrn = c.cg_calc_ur(alpha)
            beta = rrn / rro
            c.cg_calc_p(beta)
            c.local_halos(1, ["u", "p"])
            scalars += [pw, rrn]
            rro = rrn
        pw = c.cg_calc_w()
        scalars.append(pw)
        out = {n: c.read(n) for n in ("u", "p", "r", "w")}
        ratio = c.kernel_stats().get("cg_ra_finish", (0, 0.0))[0]
    return np.array(scalars), out, ratio


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("cells", [(250, 250), (1030, 300), (124, 6)])
def test_run_ahead_is_the_plain_iteration_bit_for_bit(cells):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    s0, f0, r0 = _cg(x, y, 3, False, 40)
    s1, f1, r1 = _cg(x, y, 3, True, 40)
    assert r0 == 0
    # Two run-ahead finishers per iteration.
    assert r1 >= 2 * 36, r1
    assert _same(s0, s1)
    for n in f0:
        assert _same(f0[n], f1[n]), n


def test_run_ahead_with_odd_width_stays_off():
    """x odd: the pair-per-thread kernels (and the run-ahead) do not apply."""
    x, y = 121 + 2 * H, 50 + 2 * H
    s0, f0, r0 = _cg(x, y, 5, False, 10)
    s1, f1, r1 = _cg(x, y, 5, True, 10)
    assert r0 == 0 and r1 == 0
    assert _same(s0, s1)
    for n in f0:
        assert _same(f0[n], f1[n]), n


def test_run_ahead_is_dropped_when_the_host_strays():
    """Another alpha than the device's (iteration 5), a field read (7), a
    2-norm (9) and a field summary (11) between the calls: each drops the
    run-ahead, and every result is that of the plain iteration given the
    same calls."""
    x, y = 200 + 2 * H, 120 + 2 * H

    def stray(c, it, alpha):
        if it == 5:
            return alpha * (1.0 + 2.0 ** -40)
        if it == 7:
            c.read("r")
        if it == 9:
            c.calculate_2norm("r")
        if it == 11:
            c.field_summary()
        return alpha

    s0, f0, r0 = _cg(x, y, 8, False, 20, stray)
    s1, f1, r1 = _cg(x, y, 8, True, 20, stray)
    assert r0 == 0 and r1 > 0
    assert _same(s0, s1)
    for n in f0:
        assert _same(f0[n], f1[n]), n


def test_run_ahead_whole_solves_match(tmp_path):
    """Three timesteps of a rectangle deck through the host program with
    the run-ahead on and off: the same iterations, summaries and final
    field, bit for bit, and the run-ahead taken over most passes."""
    deck = Deck.with_random_rectangles(
        512, 512, 16, problem_seed(2, 7), density="uniform", end_step=3,
        coefficient="inverse", extent=10.0)
    runs = {}
    for name, env in (("off", {"TEA_HIP_RUN_AHEAD": "0"}), ("on", {})):
        wd = tmp_path / name
        wd.mkdir()
        env = dict(env, TEA_HIP_STATS=str(wd / "stats.json"))
        rec, _ = run_deck(deck, workdir=str(wd), env=env,
                          extra_args=["--field-out", "u.bin"])
        stats = json.load(open(wd / "stats.json"))
        runs[name] = (rec, np.fromfile(wd / "u.bin"), stats)
    (a, ua, sa), (b, ub, sb) = runs["off"], runs["on"]
    assert [s["iterations"] for s in a["steps"]] == [s["iterations"] for s in b["steps"]]
    assert a["summaries"] == b["summaries"]
    assert _same(ua, ub)
    assert "cg_ra_finish" not in sa or sa["cg_ra_finish"]["launches"] == 0
    iters = sum(s["executed"] for s in b["steps"])
    assert sb["cg_ra_finish"]["launches"] >= iters, (sb["cg_ra_finish"], iters)
