"""CPU model of the stacked workgroups of the stencil passes (stack_factor,
stacked_grid and stacked_block in exploringsycl_amd/csrc/tea_hip_impl.h):
with S = x / 512 blocks per row, workgroup w of KS blocks takes blocks
c + S t, c = w % S, t = KS (w / S) + s for its s-th block.  Checked for the
grids of the tests and the benchmark: every 512-cell block is taken by
exactly one (workgroup, s), the blocks past the grid are the only ones
skipped, the blocks of a workgroup are one block row apart, and at the
benchmark's 4100-cell rows the row above a block's cells is, but for its
last 4 cells, the cells of the next block of its workgroup (the reuse the
mapping is for).  The GPU tests check the kernels' sums bit for bit."""

import os
import re

import pytest

from tests.oracle_ctypes import REPO

HDR = os.path.join(REPO, "exploringsycl_amd", "csrc", "tea_hip_impl.h")
BLOCK = 512  # cells per block (kPairBlock)


def _header_uses_this_mapping():
    text = open(HDR).read()
    body = text[text.index("__device__ inline int stacked_block"):]
    body = body[:body.index("}\n")]
    return ("blockIdx.x % S" in body and "(blockIdx.x / S) * KS" in body
            and "threadIdx.x / T" in body)


def _grid(x, y, ks):
    n = x * y
    nblk = (n + BLOCK - 1) // BLOCK
    S = x // BLOCK
    if ks == 1 or S < 1:
        return nblk, S, nblk, [(w, 0, w) for w in range(nblk)]
    tn = (nblk + S - 1) // S
    nwg = S * ((tn + ks - 1) // ks)
    taken = []
    for w in range(nwg):
        for s in range(ks):
            taken.append((w, s, w % S + S * ((w // S) * ks + s)))
    return nblk, S, nwg, taken


def test_header_mapping_is_the_modelled_one():
    assert _header_uses_this_mapping()


@pytest.mark.parametrize("cells", [(512, 61), (596, 70), (1024, 33), (2048, 19), (510, 61),
                                   (1022, 40), (1030, 300), (4096, 4096), (8192, 64)])
@pytest.mark.parametrize("ks", [1, 2, 4])
def test_every_block_once(cells, ks):
    x, y = cells[0] + 4, cells[1] + 4
    nblk, S, nwg, taken = _grid(x, y, ks)
    blocks = [b for _, _, b in taken if b < nblk]
    assert sorted(blocks) == list(range(nblk))
    # Skipped entries are past the grid, in the last row of workgroups.
    skipped = [(w, b) for w, _, b in taken if b >= nblk]
    if ks > 1:
        assert all(w >= nwg - S for w, _ in skipped)
        # The blocks of a workgroup are one block row (S blocks) apart.
        by_wg = {}
        for w, s, b in taken:
            by_wg.setdefault(w, []).append(b)
        assert all(bs[i + 1] - bs[i] == S for bs in by_wg.values() for i in range(len(bs) - 1))


def test_rows_above_come_from_the_workgroup_at_4100():
    x = 4100
    S, o = x // BLOCK, x % BLOCK
    assert (S, o) == (8, 4)
    b = 5 * S + 3  # any block with a next block row
    cells = range(BLOCK * b, BLOCK * (b + 1))
    above = [e + x for e in cells]
    nxt = range(BLOCK * (b + S), BLOCK * (b + S + 1))
    inside = [a for a in above if a in nxt]
    assert len(inside) == BLOCK - o
