"""The CPU baseline (oracle/ckernels_omp.c, the CG path in the style of the
reference's OpenMP c_kernels) solves the same problems as the oracle: the
tea.problems known answers, and the oracle host's iteration counts and
temperature on a 1/density deck (OpenMP reductions and contracted products
round differently from the oracle's tree, so counts may move by one)."""

import os

import pytest

from exploringsycl_amd.deck import Deck, config_deck
from exploringsycl_amd.runner import run_deck
from tests.oracle_ctypes import ORACLE_TEALEAF

import bench

LIB = os.path.join(bench.REPO, "oracle", "build", "libckernels_omp_v3.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="baseline not built")


@pytest.mark.parametrize("cells,steps,want", [(10, 10, 157.550841832793),
                                              (64, 1, 108.4697720003111),
                                              (250, 10, 106.272211786466)])
def test_known_answers(cells, steps, want):
    lib, _ = bench._omp_baseline_lib()
    _, iters, temp = bench._omp_run(lib, Deck(cells, cells, end_step=steps), 1)
    assert len(iters) == steps
    assert abs(temp / want - 1) < 1e-9


def test_matches_the_oracle_host(tmp_path):
    deck = config_deck(2, x_cells=96, end_step=3)
    rec, _ = run_deck(deck, binary=ORACLE_TEALEAF, workdir=str(tmp_path))
    lib, _ = bench._omp_baseline_lib()
    _, iters, temp = bench._omp_run(lib, deck, 2)
    want = [s["executed"] for s in rec["steps"]]
    assert all(abs(a - b) <= 1 for a, b in zip(iters, want)), (iters, want)
    assert abs(temp / rec["summaries"][-1]["temp"] - 1) < 1e-9
