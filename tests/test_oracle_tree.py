"""Pins the oracle's reduction to the reference's tree
(TeaLeaf/c_kernels/sycl/sycl_shared.hpp:35-79, cg.cpp:177-183): 64-wide
pairwise trees over zero-padded groups, repeated over the partials."""

import numpy as np


def test_pairwise_not_sequential(oracle):
    # Worked by hand: the tree adds (1 + 1e16) and (-1e16 + 1) first; both
    # round to +-1e16 (1e16 + 1 is halfway between representable neighbours
    # 1e16 and 1e16 + 2 and rounds to the even 1e16), so the sum is 0.
    # Left to right it would be ((1 + 1e16) - 1e16) + 1 = 1.
    v = np.zeros(64)
    v[:4] = [1.0, 1e16, -1e16, 1.0]
    assert oracle.tree_sum(v) == 0.0
    assert ((1.0 + 1e16) - 1e16) + 1.0 == 1.0


def test_second_level_is_a_tree_over_partials(oracle):
    # Group sums 1e16, 0, -1e16, 1: the level-2 tree adds (1e16 + 0) and
    # (-1e16 + 1) = -1e16 (rounded) first, giving 0; summing the groups left
    # to right would give (1e16 - 1e16) + 1 = 1.
    v = np.zeros(4 * 64)
    v[0] = 1e16
    v[2 * 64] = -1e16
    v[3 * 64] = 1.0
    assert oracle.tree_sum(v) == 0.0


def test_exact_for_integers(oracle):
    rng = np.random.default_rng(0)
    for n in (1, 63, 64, 65, 4095, 4096, 4097, 262145, 300001):
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        assert oracle.tree_sum(v) == float(v.sum()), n


def test_padding_is_zero(oracle):
    v = np.arange(1, 101, dtype=np.float64)
    assert oracle.tree_sum(v) == 5050.0
    assert oracle.tree_sum(np.array([2.5])) == 2.5


def test_three_levels(oracle):
    # 64**2 * 3 + 5 elements need three passes over the partials.
    n = 64 * 64 * 3 + 5
    v = np.full(n, 0.5)
    assert oracle.tree_sum(v) == 0.5 * n
