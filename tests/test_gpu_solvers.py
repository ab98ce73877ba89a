"""Jacobi, Chebyshev and PPCG on the GPU (libteahip.so through its C ABI and
bin/tealeaf) against the CPU oracle: every field and every returned scalar
bitwise equal, in the fused one-pass kernels (even rows), the one-launch-
per-reference-kernel mode and on odd rows; whole solves give the oracle
host's iteration counts, eigenvalue bounds and field summaries, and the
reference's own host driving the HIP backend prints the reference host's
numbers."""

import json
import os
import re
import subprocess

import numpy as np
import pytest

from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import TEA_PROBLEMS, run_deck
from tests.oracle_ctypes import ORACLE_TEALEAF, REFHOST_HIP
from tests.test_gpu_kernels import H, _assert_fields, _random_fields, _same, _upload

pytestmark = pytest.mark.gpu

# (x_cells, y_cells): odd rows with one group, even rows with one group (the
# pair kernels without a reduction), a short and wide chunk, ragged spans,
# odd rows over many spans, several spans, and the benchmark size; then, for
# the two-step passes' 16 x 64 tiles, exactly one tile, one tile and one
# cell pair more each way, and two tiles and one pair each way.
SIZES = [(1, 1), (2, 2), (60, 3), (250, 250), (601, 599), (1030, 300),
         (4096, 4096), (64, 16), (66, 18), (130, 34)]
# Fused (the default), and one launch per reference kernel.
MODES = ["fused", "unfused"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _chunk(x, y, mode):
    from exploringsycl_amd.chunk import DeviceChunk
    c = DeviceChunk(x, y, H)
    if mode == "unfused":
        c.set_fused(False)
    return c


def _dims(cells):
    return cells[0] + 2 * H, cells[1] + 2 * H


@pytest.mark.parametrize("cells", SIZES)
@pytest.mark.parametrize("coef", [1, 2])
def test_jacobi_init(oracle, cells, coef):
    x, y = _dims(cells)
    f = _random_fields(x, y, 21)
    names = ["density", "energy", "u", "u0", "kx", "ky"]
    with _chunk(x, y, "fused") as c:
        _upload(c, f, names)
        c.jacobi_init(coef, 0.37, 0.61)
        oracle.jacobi_init(x, y, H, coef, 0.37, 0.61, f)
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", SIZES)
@pytest.mark.parametrize("mode", MODES)
def test_jacobi_iterate(oracle, cells, mode):
    x, y = _dims(cells)
    f = _random_fields(x, y, 22)
    names = ["u", "u0", "r", "kx", "ky"]
    with _chunk(x, y, mode) as c:
        _upload(c, f, names)
        for it in range(3):
            got = c.jacobi_iterate()
            want = oracle.jacobi_iterate(x, y, H, f)
            assert _same(got, want), (it, got, want)
            c.local_halos(1, ["u"])
            oracle.halo_field(x, y, H, 1, f["u"])
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", SIZES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("steps", [1, 2, 3, 10])
def test_cheby(oracle, cells, mode, steps):
    """cheby_init, then `steps` Chebyshev steps each followed by the depth-1
    U halo update, as the Chebyshev driver calls them.  In the fused mode
    consecutive steps run two to a pass."""
    x, y = _dims(cells)
    f = _random_fields(x, y, 23)
    names = ["u", "u0", "p", "r", "w", "kx", "ky"]
    rng = np.random.default_rng(100 + steps)
    coefs = list(zip(rng.uniform(-0.5, 1.0, steps), rng.uniform(-0.5, 1.5, steps)))
    with _chunk(x, y, mode) as c:
        _upload(c, f, names)
        c.cheby_init(1.75)
        oracle.cheby_init(x, y, H, 1.75, f)
        _assert_fields(c, f, names)
        for a, b in coefs:
            c.cheby_iterate(a, b)
            oracle.cheby_iterate(x, y, H, a, b, f)
            c.local_halos(1, ["u"])
            oracle.halo_field(x, y, H, 1, f["u"])
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", [(250, 250), (601, 599)])
def test_cheby_held_step_flushed_by_other_calls(oracle, cells):
    """A held-back Chebyshev step runs before a 2-norm of r, before a halo
    update of more than U, and before a step that follows without a halo
    update; a held-back U halo update runs with it."""
    x, y = _dims(cells)
    f = _random_fields(x, y, 26)
    names = ["u", "u0", "p", "r", "w", "kx", "ky"]
    with _chunk(x, y, "fused") as c:
        _upload(c, f, names)
        c.cheby_iterate(0.3, 0.6)
        oracle.cheby_iterate(x, y, H, 0.3, 0.6, f)
        assert _same(c.calculate_2norm("r"), oracle.calculate_2norm(x, y, H, f["r"]))
        c.cheby_iterate(0.4, 0.2)
        oracle.cheby_iterate(x, y, H, 0.4, 0.2, f)
        c.local_halos(1, ["u", "p"])
        oracle.halo_field(x, y, H, 1, f["u"])
        oracle.halo_field(x, y, H, 1, f["p"])
        c.cheby_iterate(0.5, 0.1)
        oracle.cheby_iterate(x, y, H, 0.5, 0.1, f)
        c.cheby_iterate(0.6, 0.3)
        oracle.cheby_iterate(x, y, H, 0.6, 0.3, f)
        c.local_halos(1, ["u"])
        oracle.halo_field(x, y, H, 1, f["u"])
        assert _same(c.calculate_2norm("u"), oracle.calculate_2norm(x, y, H, f["u"]))
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", SIZES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("steps", [1, 2, 3, 10])
def test_ppcg(oracle, cells, mode, steps):
    """ppcg_init, then `steps` inner steps each after a depth-1 SD halo
    update, as the PPCG driver calls them.  In the fused mode consecutive
    steps run two to a pass; an odd last step runs alone when the fields are
    read back."""
    x, y = _dims(cells)
    f = _random_fields(x, y, 24)
    names = ["sd", "r", "u", "kx", "ky"]
    rng = np.random.default_rng(steps)
    coefs = list(zip(rng.uniform(-0.5, 1.0, steps), rng.uniform(-0.5, 1.5, steps)))
    with _chunk(x, y, mode) as c:
        _upload(c, f, names)
        c.ppcg_init(2.5)
        oracle.ppcg_init(x, y, H, 2.5, f)
        _assert_fields(c, f, names)
        for a, b in coefs:
            c.local_halos(1, ["sd"])
            oracle.halo_field(x, y, H, 1, f["sd"])
            c.ppcg_inner_iteration(a, b)
            oracle.ppcg_inner_iteration(x, y, H, a, b, f)
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", [(250, 250), (601, 599)])
def test_ppcg_held_step_flushed_by_other_calls(oracle, cells):
    """A held-back inner step (and its held-back SD halo update) is run
    before any other entry point: a 2-norm of sd (named by its pre-step
    address), another halo update, a second step without a halo between."""
    x, y = _dims(cells)
    f = _random_fields(x, y, 25)
    names = ["sd", "r", "u", "kx", "ky"]
    with _chunk(x, y, "fused") as c:
        _upload(c, f, names)
        c.ppcg_inner_iteration(0.3, 0.6)
        oracle.ppcg_inner_iteration(x, y, H, 0.3, 0.6, f)
        assert _same(c.calculate_2norm("sd"), oracle.calculate_2norm(x, y, H, f["sd"]))
        c.ppcg_inner_iteration(0.4, 0.2)
        oracle.ppcg_inner_iteration(x, y, H, 0.4, 0.2, f)
        c.local_halos(1, ["sd", "u"])
        oracle.halo_field(x, y, H, 1, f["sd"])
        oracle.halo_field(x, y, H, 1, f["u"])
        c.ppcg_inner_iteration(0.5, 0.1)
        oracle.ppcg_inner_iteration(x, y, H, 0.5, 0.1, f)
        c.ppcg_inner_iteration(0.6, 0.3)
        oracle.ppcg_inner_iteration(x, y, H, 0.6, 0.3, f)
        c.local_halos(1, ["sd"])
        oracle.halo_field(x, y, H, 1, f["sd"])
        assert _same(c.calculate_2norm("r"), oracle.calculate_2norm(x, y, H, f["r"]))
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", [(250, 250), (601, 599), (66, 18)])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("steps,interrupt", [(1, False), (2, False), (3, False),
                                             (10, False), (10, True)])
def test_ppcg_outer_steps_with_deferred_cg_updates(oracle, cells, mode, steps,
                                                   interrupt):
    """PPCG outer steps as the driver makes them (ppcg_driver.c:87-189):
    w = A p, u += alpha p with r -= alpha w, sd = r / theta, the inner steps
    each after an SD halo update, the 2-norm of r, p = beta p + r and the P
    halo update.  In the fused mode u += alpha p is applied by the first
    two-step pass and p = beta p + r with its ring by the next w = A p; with
    `interrupt`, a 2-norm of u after the first inner step must see u += alpha
    p applied before that step."""
    x, y = _dims(cells)
    f = _random_fields(x, y, 26)
    names = ["p", "w", "u", "r", "sd", "kx", "ky"]
    rng = np.random.default_rng(100 + steps)
    coefs = list(zip(rng.uniform(-0.5, 1.0, steps), rng.uniform(-0.5, 1.5, steps)))
    with _chunk(x, y, mode) as c:
        _upload(c, f, names)
        for outer, (alpha, beta) in enumerate([(0.7, 0.4), (0.35, 0.9)]):
            assert _same(c.cg_calc_w(), oracle.cg_calc_w(x, y, H, f))
            assert _same(c.cg_calc_ur(alpha), oracle.cg_calc_ur(x, y, H, alpha, f))
            c.ppcg_init(2.5)
            oracle.ppcg_init(x, y, H, 2.5, f)
            for k, (a, b) in enumerate(coefs):
                c.local_halos(1, ["sd"])
                oracle.halo_field(x, y, H, 1, f["sd"])
                c.ppcg_inner_iteration(a, b)
                oracle.ppcg_inner_iteration(x, y, H, a, b, f)
                if interrupt and k == 0 and outer == 0:
                    assert _same(c.calculate_2norm("u"),
                                 oracle.calculate_2norm(x, y, H, f["u"]))
            assert _same(c.calculate_2norm("r"), oracle.calculate_2norm(x, y, H, f["r"]))
            c.cg_calc_p(beta)
            oracle.cg_calc_p(x, y, H, beta, f)
            c.local_halos(1, ["p"])
            oracle.halo_field(x, y, H, 1, f["p"])
        assert _same(c.cg_calc_w(), oracle.cg_calc_w(x, y, H, f))
        _assert_fields(c, f, names)


def test_field_addresses_follow_the_exchange():
    """The fused passes exchange buffers; tea_hip_field reports the current
    ones, and the 2-norm entry point accepts them."""
    with _chunk(64 + 2 * H, 32 + 2 * H, "fused") as c:
        u_before, _ = c.field_info("u")
        r_before, _ = c.field_info("r")
        c.jacobi_iterate()
        assert c.field_info("u")[0] == r_before
        assert c.field_info("r")[0] == u_before
        c.cheby_iterate(0.5, 0.5)  # held back for a possible pair
        c.calculate_2norm("u")  # runs it; u is still named by its old address
        assert c.field_info("u")[0] not in (u_before, r_before)
        c.calculate_2norm("r")
        sd_before, _ = c.field_info("sd")
        r_before, _ = c.field_info("r")
        c.ppcg_inner_iteration(0.5, 0.5)  # held back for a possible pair
        c.calculate_2norm("sd")  # runs it; sd is still named by its old address
        assert c.field_info("sd")[0] != sd_before
        c.local_halos(1, ["sd"])
        c.ppcg_inner_iteration(0.5, 0.5)
        c.local_halos(1, ["sd"])
        c.ppcg_inner_iteration(0.5, 0.5)  # with the one before: one pass
        c.calculate_2norm("r")
        assert c.field_info("r")[0] != r_before


# ---- whole solves ------------------------------------------------------------------

DECKS = {
    "jacobi_64x64x1": Deck(64, 64, end_step=1, extra=["use_jacobi"]),
    "jacobi_odd_45x38x2": Deck.with_random_rectangles(
        45, 38, 6, problem_seed(9, 3), end_step=2, extra=["use_jacobi", "eps 1e-12"]),
    "cheby_250x250x2": Deck(250, 250, end_step=2, extra=["use_chebyshev"]),
    "cheby_recip_200x300x2": Deck.with_random_rectangles(
        200, 300, 16, problem_seed(2), end_step=2, coefficient="inverse",
        extra=["use_chebyshev", "presteps 40"]),
    "cheby_odd_99x77x2": Deck.with_random_rectangles(
        99, 77, 8, problem_seed(2, 5), end_step=2, coefficient="inverse",
        extra=["use_chebyshev", "errswitch", "epslim 1e-4"]),
    "ppcg_250x250x2": Deck(250, 250, end_step=2, extra=["use_ppcg"]),
    "ppcg_recip_200x300x2": Deck.with_random_rectangles(
        200, 300, 16, problem_seed(2), end_step=2, coefficient="inverse",
        extra=["use_ppcg", "ppcg_inner_steps 6"]),
    "ppcg_odd_99x77x2": Deck.with_random_rectangles(
        99, 77, 8, problem_seed(2, 5), end_step=2, coefficient="inverse",
        extra=["use_ppcg", "errswitch"]),
}
HOST_MODES = {"fused_pairs": {"TEA_HIP_FUSED": "1"},
              "fused_single": {"TEA_HIP_FUSED": "1", "TEA_HIP_PAIRS": "0"},
              "unfused": {"TEA_HIP_FUSED": "0"}}


@pytest.mark.parametrize("name", sorted(DECKS))
@pytest.mark.parametrize("mode", sorted(HOST_MODES))
def test_solves_match_oracle_host(name, mode):
    deck = DECKS[name]
    gpu, _ = run_deck(deck, env=HOST_MODES[mode])
    cpu, _ = run_deck(deck, binary=ORACLE_TEALEAF, env={"OMP_NUM_THREADS": "16"})
    for key in ("iterations", "solver_iterations", "estimated", "eigmin", "eigmax"):
        assert [s[key] for s in gpu["steps"]] == [s[key] for s in cpu["steps"]], key
    assert len(gpu["summaries"]) == len(cpu["summaries"])
    for a, b in zip(gpu["summaries"], cpu["summaries"]):
        for k in ("vol", "mass", "ie", "temp"):
            assert a[k] == b[k], (k, a, b)


@pytest.mark.parametrize("solver", ["use_chebyshev", "use_ppcg"])
@pytest.mark.parametrize("cells,steps", [(512, 20), (1024, 20), (4096, 1)])
def test_known_answers_with_each_solver(solver, cells, steps):
    """TeaLeaf/tea.problems rows solved with Chebyshev and PPCG: the host's
    1e-5 check, and 1e-6 relative (these solvers stop when the squared
    residual is below eps, as the reference's drivers do)."""
    rec, _ = run_deck(Deck(cells, cells, end_step=steps, extra=[solver]), timeout=900)
    assert rec["check_found"] and rec["check_passed"]
    temp = rec["summaries"][-1]["temp"]
    assert abs(temp / rec["check_value"] - 1.0) < 1e-6, (temp, rec["check_value"])


def _refhost_fixtures():
    with open(os.path.join(GOLDEN, "refhost_oracle.json")) as f:
        return json.load(f)


@pytest.mark.skipif(not os.path.exists(REFHOST_HIP),
                    reason="reference host build needs the reference tree")
@pytest.mark.parametrize("name", sorted(n for n in _refhost_fixtures()
                                        if n.split("_")[0] in ("jacobi", "cheby", "ppcg")))
def test_reference_host_drives_hip_solvers(tmp_path, name):
    """The reference's own host (compiled unchanged) on the HIP backend prints
    what it prints on the oracle backend: iteration counts, Chebyshev/PPCG
    counts, eigenvalue bounds and the final temperature."""
    want = _refhost_fixtures()[name]
    (tmp_path / "tea.in").write_text(want["deck"])
    (tmp_path / "tea.problems").write_text(open(TEA_PROBLEMS).read())
    p = subprocess.run([REFHOST_HIP], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    lines = p.stdout.splitlines()
    iters = [int(l.split()[1]) for l in lines
             if l.startswith("CG:") or l.startswith("Jacobi:")]
    assert iters == want["iterations"]
    after = []
    for l in lines:
        m = re.match(r"(Cheby|PPCG):\s+(\d+) iterations(?: \((\d+) estimated\))?", l)
        if m:
            after.append([int(m.group(2)), int(m.group(3)) if m.group(3) else -1])
    assert after == want["after_switch"]
    mins = [l.split()[2] for l in lines if l.startswith("Min. eigenvalue")]
    maxs = [l.split()[2] for l in lines if l.startswith("Max. eigenvalue")]
    assert [list(e) for e in zip(mins, maxs)] == want["eigenvalues"]
    actual = [l.split()[1] for l in lines if l.startswith("Actual")]
    assert actual[-1] == want["final_temp_printed"]
