"""The mixed-precision u update held over two w-steps: with four cells per
lane two mpcg_calc_w / fcg_calc_w calls in turn hold their u += alpha p
back and the third makes all three in one pass over u.  These
tests run the drivers' call sequences with the hold on and off
(TEA_HIP_HOLD_U=0) and require the same scalars and fields bit for bit, also
when other calls come between the steps (reads of u, refreshes after an odd
and an even number of steps, a 2-norm, a field summary), each of which must
make the held update first."""

import os

import numpy as np
import pytest

from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import run_deck

pytestmark = pytest.mark.gpu

H = 2


def _chunk(x, y, hold):
    from exploringsycl_amd.chunk import DeviceChunk
    old = os.environ.get("TEA_HIP_HOLD_U")
    os.environ["TEA_HIP_HOLD_U"] = "1" if hold else "0"
    try:
        return DeviceChunk(x, y, H)
    finally:
        if old is None:
            del os.environ["TEA_HIP_HOLD_U"]
        else:
            os.environ["TEA_HIP_HOLD_U"] = old


def _fields(x, y, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.1, 100.0, x * y), rng.uniform(0.1, 25.0, x * y)


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _mpcg(x, y, seed, hold, iters, refresh_at=(), stray=None):
    """mpcg_init, then the mixed-precision driver's sequence (flexible
    recursion); a refresh after the iterations in `refresh_at`; `stray(c,
    it)` may act between the calls.  Returns the scalars, the final u and
    the number of separate u-update launches."""
    density, energy = _fields(x, y, seed)
    with _chunk(x, y, hold) as c:
        c.timing(True)
        c.write("density", density)
        c.write("energy", energy)
        rr, _ = c.mpcg_init(2, 0.8, 1.3)
        scalars = [rr]
        alpha = beta = 0.0
        pu = pp = False
        for it in range(iters):
            pw, pr = c.mpcg_calc_w(alpha, pu, beta, pp)
            alpha = pr / pw
            if stray is not None:
                stray(c, it)
            rrn, rw = c.mpcg_calc_r(alpha)
            pu = True
            scalars += [pw, pr, rrn, rw]
            if it in refresh_at:
                a, b, rw = c.mpcg_refresh(alpha, True)
                scalars += [a, b, rw]
                pu = False
            beta = -rw / pw
            pp = True
        if pu:
            scalars += list(c.mpcg_refresh(alpha, True))
        u = c.read("u")
        applies = c.kernel_stats().get("mpcg_apply_u", (0, 0.0))[0]
    return np.array(scalars), u, applies


@pytest.mark.parametrize("cells", [(124, 60), (252, 131), (121, 50)])
def test_held_u_is_the_step_by_step_update_bit_for_bit(cells):
    """25 steps with refreshes after steps 6 and 12 and at the end; x = 128
    and 256 take the four-cell kernels (and the hold), x = 125 not.  Of the
    steps with a u update, two hold theirs and the third makes all three:
    the refresh after step 6 finds none held, those after 12 and 24 two
    each, each made by a launch of its own."""
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    s0, u0, n0 = _mpcg(x, y, 3, False, 25, refresh_at=(6, 12))
    s1, u1, n1 = _mpcg(x, y, 3, True, 25, refresh_at=(6, 12))
    assert _same(s0, s1)
    assert _same(u0, u1)
    assert n0 == 3
    assert n1 == (n0 + 4 if x % 4 == 0 else n0), (n0, n1)


def test_held_u_is_made_before_other_calls_read_u():
    """A read of u (iteration 4), a 2-norm of u (7), a field summary (10)
    and a U halo update (13) between the steps: each sees u with every
    reported update made."""
    x, y = 200 + 2 * H, 90 + 2 * H
    seen = {}

    def stray(tag):
        def f(c, it):
            if it == 4:
                seen.setdefault(tag, []).append(c.read("u"))
            if it == 7:
                seen.setdefault(tag, []).append(c.calculate_2norm("u"))
            if it == 10:
                seen.setdefault(tag, []).append(np.array(c.field_summary()))
            if it == 13:
                c.local_halos(1, ["u"])
        return f

    s0, u0, _ = _mpcg(x, y, 9, False, 16, stray=stray("off"))
    s1, u1, _ = _mpcg(x, y, 9, True, 16, stray=stray("on"))
    assert _same(s0, s1)
    assert _same(u0, u1)
    for a, b in zip(seen["off"], seen["on"]):
        assert _same(a, b)


def _fcg(x, y, seed, hold, iters):
    """Flexible CG with the Jacobi preconditioner through the FCG entry
    points (fcg_calc_w holds u as mpcg_calc_w does), as the FCG driver calls
    them, with a refresh after iteration 8."""
    density, energy = _fields(x, y, seed)
    with _chunk(x, y, hold) as c:
        c.write("density", density)
        c.write("energy", energy)
        c.fcg_attach()
        rr, _ = c.fcg_init(2, 0.8, 1.3)
        zw = c.fcg_precondition(False)
        scalars = [rr, zw]
        alpha = beta = 0.0
        pu = False
        for it in range(iters):
            pw, pr = c.fcg_calc_w(alpha, pu, beta)
            alpha = pr / pw
            rrn, zw = c.fcg_calc_r_precondition(alpha, False)
            pu = True
            scalars += [pw, pr, rrn, zw]
            if it == 8:
                a, b, _ = c.mpcg_refresh(alpha, True)
                zw = c.fcg_precondition(False)
                scalars += [a, b, zw]
                pu = False
            beta = -zw / pw
        if pu:
            scalars += list(c.mpcg_refresh(alpha, True))
        return np.array(scalars), c.read("u"), c.read_z()


def test_held_u_in_flexible_cg():
    x, y = 176 + 2 * H, 100 + 2 * H
    s0, u0, z0 = _fcg(x, y, 5, False, 15)
    s1, u1, z1 = _fcg(x, y, 5, True, 15)
    assert _same(s0, s1)
    assert _same(u0, u1)
    assert np.array_equal(z0.view(np.int32), z1.view(np.int32))


@pytest.mark.parametrize("solver", ["mpcg", "fcgs"])
def test_held_u_whole_solves_match(tmp_path, solver):
    """Two timesteps of a 1024^2 rectangle deck through the host program with
    the hold on and off: the same iterations, summaries and final field, bit
    for bit (refresh interval 25: refreshes after odd and even step
    counts)."""
    extra = {"mpcg": ["use_mpcg"], "fcgs": ["use_fcg", "fcg_coarse_sweeps=6"]}[solver]
    deck = Deck.with_random_rectangles(
        1024, 1024, 24, problem_seed(5, 11), density="log10", density_range=(-2.0, 2.0),
        end_step=2, coefficient="inverse", extent=2.5,
        extra=["rel_eps=1e-10", "refresh_interval=25"] + extra)
    runs = {}
    for name, env in (("off", {"TEA_HIP_HOLD_U": "0"}), ("on", {})):
        wd = tmp_path / name
        wd.mkdir()
        rec, _ = run_deck(deck, workdir=str(wd), env=env, extra_args=["--field-out", "u.bin"])
        runs[name] = (rec, np.fromfile(wd / "u.bin"))
    (a, ua), (b, ub) = runs["off"], runs["on"]
    assert [s["iterations"] for s in a["steps"]] == [s["iterations"] for s in b["steps"]]
    assert [s["refreshes"] for s in a["steps"]] == [s["refreshes"] for s in b["steps"]]
    assert a["summaries"] == b["summaries"]
    assert _same(ua, ub)
