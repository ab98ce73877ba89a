"""Pins the oracle's flexible CG with the Jacobi preconditioner: with a
fixed SPD preconditioner FCG(1) is preconditioned CG, so it must reach the
FP64 CG solution (itself pinned by a dense solve) to the stopping tolerance,
and its recursive residual must agree with the FP64 true residual."""

import os

import pytest

from exploringsycl_amd.deck import Deck, problem_seed
from exploringsycl_amd.runner import run_deck
from tests.oracle_ctypes import ORACLE_TEALEAF

pytestmark = pytest.mark.skipif(not os.path.exists(ORACLE_TEALEAF),
                                reason="oracle host not built")

BASE = dict(density="log10", density_range=(-2.0, 2.0), end_step=2,
            coefficient="inverse")


@pytest.mark.parametrize("cells,seed", [(10, 0), (64, 1), (96, 2)])
def test_jacobi_pcg_reaches_the_cg_solution(tmp_path, cells, seed):
    def deck(extra):
        return Deck.with_random_rectangles(cells, cells, 12, problem_seed(5, seed),
                                           extra=extra, **BASE)
    cg, _ = run_deck(deck(["rel_eps=1e-13"]), binary=ORACLE_TEALEAF,
                     workdir=str(tmp_path / "cg"))
    jp, _ = run_deck(deck(["use_jacobi_pcg", "rel_eps=1e-10"]), binary=ORACLE_TEALEAF,
                     workdir=str(tmp_path / "jpcg"))
    assert jp["solver"] == "Jacobi-PCG"
    for s in jp["steps"]:
        assert s["rel_residual"] < 1e-10
        assert s["fallback_at"] == -1
    a, b = cg["summaries"][-1], jp["summaries"][-1]
    for k in ("vol", "mass", "ie"):
        assert a[k] == b[k]
    # |u - u*| <= |A^-1| |r|; with |A^-1| <= 1 (A = I + SPD) and
    # |r| <= 1e-10 |b| the temperature moves by far less than 1e-8.
    assert abs(a["temp"] / b["temp"] - 1) < 1e-8


def test_learned_solver_needs_a_network(tmp_path):
    d = Deck(16, 16, end_step=1, extra=["use_fcg", "rel_eps=1e-10"])
    with pytest.raises(RuntimeError):
        run_deck(d, binary=ORACLE_TEALEAF, workdir=str(tmp_path))
