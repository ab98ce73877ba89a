"""The benchmark configurations at full size on the GPU.

- The bench workload (4096^2, K = 1/density, tea.in + 16 seeded rectangles,
  eps 1e-15, max 10000 iterations), first timestep: iteration count, the
  four field summaries and the final field bit for bit against the CPU
  oracle's run of the same deck (tests/golden/config2_4096_oracle.json,
  written by tests/golden/make_golden.py config2), in the fused and the
  one-launch-per-reference-kernel modes.  The reference's own check of this
  quantity is TeaLeaf/drivers/field_summary_driver.c:8-52.
- Mixed-precision CG on the mixed-precision workload (4096^2, 32 seeded
  rectangles, log10 density uniform in [-1, 2], FP64 refresh every 50
  iterations, 1e-10 relative residual) against plain FP64 CG to the same
  tolerance.
- Flexible CG with the learned preconditioner at 8192^2 (64 rectangles,
  log10 density in [-2, 2], an unseen seed) against mixed-precision CG.
"""

import hashlib
import json
import os
import shutil

import numpy as np
import pytest

from exploringsycl_amd import learned
from exploringsycl_amd.deck import Deck, config_deck, problem_seed
from exploringsycl_amd.runner import run_deck

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "config2_4096_oracle.json")


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="fixture not written")
@pytest.mark.parametrize("fused", ["1", "0"])
def test_bench_workload_first_step_matches_oracle_fixture(tmp_path, fused):
    fx = json.load(open(FIXTURE))
    deck = config_deck(2, x_cells=fx["cells"], end_step=fx["steps"])
    assert deck.render() == fx["deck"]
    rec, _ = run_deck(deck, workdir=str(tmp_path), env={"TEA_HIP_FUSED": fused},
                      extra_args=["--field-out", "u.bin"], timeout=900)
    assert [s["iterations"] for s in rec["steps"]] == fx["iterations"]
    assert [s["executed"] for s in rec["steps"]] == fx["executed"]
    for got, want in zip(rec["summaries"], fx["summaries"]):
        for k in ("vol", "mass", "ie", "temp"):
            assert float(got[k]).hex() == want[k], (k, got[k], float.fromhex(want[k]))
    with open(tmp_path / "u.bin", "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == fx["u_sha256"]


def _config3(extra):
    return Deck.with_random_rectangles(
        4096, 4096, 32, problem_seed(3, 0), density="log10", density_range=(-1.0, 2.0),
        end_step=1, coefficient="inverse", max_iters=20000,
        extra=["rel_eps=1e-10", "refresh_interval=50"] + extra)


def test_mixed_precision_workload_against_fp64(tmp_path):
    """FP32 stencil and vectors with FP64 dots, scalars, solution and the
    true-residual refresh reach the FP64 solver's answer: both stop at
    |b - A u| < 1e-10 |b|, so (|A^-1| <= 1) the fields differ by at most
    2e-10 |b| in the 2-norm; the summaries' relative difference is
    reported against the FP64 reference."""
    runs = {}
    for name, extra in (("fp64", []), ("mp", ["use_mpcg"])):
        rec, _ = run_deck(_config3(extra), workdir=str(tmp_path / name),
                          extra_args=["--field-out", "u.bin"], timeout=900)
        runs[name] = (rec, np.fromfile(tmp_path / name / "u.bin"))
    (a, ua), (b, ub) = runs["fp64"], runs["mp"]
    sa, sb = a["steps"][0], b["steps"][0]
    print(f"FP64 CG {sa['iterations']} iterations, mixed precision {sb['iterations']} "
          f"({sb['refreshes']} refreshes)")
    assert sa["rel_residual"] < 1e-10 and sb["rel_residual"] < 1e-10
    assert abs(sb["iterations"] - sa["iterations"]) <= 0.1 * sa["iterations"]
    for k in ("vol", "mass", "ie"):   # unaffected by the solve
        assert a["summaries"][-1][k] == b["summaries"][-1][k]
    ta, tb = a["summaries"][-1]["temp"], b["summaries"][-1]["temp"]
    print(f"temperature relative error {abs(tb / ta - 1):.3e}, field max relative "
          f"{np.abs(ub - ua).max() / np.abs(ua).max():.3e}")
    assert abs(tb / ta - 1) < 1e-9
    assert np.linalg.norm(ub - ua) <= 1e-9 * np.linalg.norm(ua)


@pytest.mark.skipif(not os.path.exists(learned.WEIGHTS_PATH), reason="no network file")
def test_learned_fcg_at_8192(tmp_path):
    """Config 5 at 8192^2: the learned flexible CG reaches 1e-10 without the
    fallback, agrees with mixed-precision CG, and needs well under a third
    of its iterations."""
    def deck(extra):
        return Deck.with_random_rectangles(
            8192, 8192, 64, problem_seed(5, 3000), density="log10",
            density_range=(-2.0, 2.0), end_step=1, coefficient="inverse",
            max_iters=60000, extra=["rel_eps=1e-10"] + extra)
    wd = tmp_path / "fcg"
    wd.mkdir()
    shutil.copy(learned.WEIGHTS_PATH, wd / "net.bin")
    fcg, _ = run_deck(deck(["use_fcg", "fcg_weights=net.bin"]), workdir=str(wd),
                      extra_args=["--field-out", "u.bin"], timeout=900)
    mp, _ = run_deck(deck(["use_mpcg"]), workdir=str(tmp_path / "mp"),
                     extra_args=["--field-out", "u.bin"], timeout=900)
    a, b = fcg["steps"][0], mp["steps"][0]
    print(f"8192^2: learned FCG {a['iterations']} iterations {a['seconds']:.2f} s, "
          f"mixed-precision CG {b['iterations']} iterations {b['seconds']:.2f} s")
    assert a["rel_residual"] < 1e-10 and b["rel_residual"] < 1e-10
    assert a["fallback_at"] == -1
    assert 3 * a["iterations"] < b["iterations"]
    ta, tb = fcg["summaries"][-1]["temp"], mp["summaries"][-1]["temp"]
    assert abs(ta / tb - 1) < 1e-8
    ua = np.fromfile(wd / "u.bin")
    ub = np.fromfile(tmp_path / "mp" / "u.bin")
    assert np.linalg.norm(ua - ub) <= 1e-9 * np.linalg.norm(ub)
