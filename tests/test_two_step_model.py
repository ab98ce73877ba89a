"""CPU model of the two-step passes' tiling (ppcg_two_step_kernel and
cheby_two_step_kernel in exploringsycl_amd/csrc/solvers.hip): the same tile
shape, staging, margin, ring-mirror and edge-ownership index arithmetic and
the same expressions, one block at a time.  Checked bitwise against two
oracle steps with the depth-1 halo update between them, with every field
access bounds-checked, every output cell written exactly once, and the
margin cells outside the interior, which the kernels leave unset, never
read.  The GPU
tests check the kernels themselves; this pins the index logic on the CPU."""

import os
import re

import numpy as np
import pytest

from tests.oracle_ctypes import REPO

SRC = os.path.join(REPO, "exploringsycl_amd", "csrc", "solvers.hip")


def _tile_shape():
    text = open(SRC).read()
    tx = int(re.search(r"constexpr int kTileX = (\d+);", text).group(1))
    ty = int(re.search(r"constexpr int kTileY = (\d+);", text).group(1))
    return tx, ty


TX, TY = _tile_shape()
S0W, S0H, S1W, S1H = TX + 4, TY + 4, TX + 2, TY + 2


def tile_of_block(b, nt):
    """Block b's tile: XCD k = b % 8 takes the k-th eighth of the tiles."""
    k, per, rem = b & 7, nt >> 3, nt & 7
    return k * per + min(k, rem) + (b >> 3)


@pytest.mark.parametrize("nt", [1, 5, 8, 9, 63, 64, 65, 1000, 4096 // 8 * 32 + 3])
def test_tile_of_block_is_a_permutation(nt):
    assert sorted(tile_of_block(b, nt) for b in range(nt)) == list(range(nt))


def reflect_staged_ring(S0, r0, c0, h, xe, ye):
    """The kernels' held-back update before step k: ring cells of the staged
    area take their mirror's value (double mirror at the corners)."""
    upd = {}
    for a in range(S0H):
        for b in range(S0W):
            jj, kc = r0 - 2 + a, c0 - 2 + b
            if not (h - 1 <= jj <= ye and h - 1 <= kc <= xe) or (h <= jj < ye and h <= kc < xe):
                continue
            mj = h if jj == h - 1 else (ye - 1 if jj == ye else jj)
            mk = h if kc == h - 1 else (xe - 1 if kc == xe else kc)
            upd[(a, b)] = S0[mj - r0 + 2][mk - c0 + 2]
    for (a, b), v in upd.items():
        S0[a][b] = v


def op5(c, l, r, b, t, kx0, kx1, ky0, ky1):
    diag = (1.0 + (kx1 + kx0)) + (ky1 + ky0)
    return (diag * c - (kx1 * r + kx0 * l)) - (ky1 * t + ky0 * b)

def model_ppcg(x, y, h, sd_old, sd_new, r, r_new, u, kx, ky, a0, b0, a1, b1,
               p_pend=None, alpha_pend=0.0, pre_halo=False):
    n = x * y
    def G(arr, i):
        assert 0 <= i < n, ("read OOB", i)
        return arr[i]
    def P(arr, i, v):
        assert 0 <= i < n, ("write OOB", i)
        arr[i] = v
    xe, ye = x - h, y - h
    tiles_x = (x - 2 * h + TX - 1) // TX
    tiles_y = (y - 2 * h + TY - 1) // TY
    written = np.zeros(n, dtype=int)
    for blk in range(tiles_x * tiles_y):
        tile = tile_of_block(blk, tiles_x * tiles_y)
        txi, tyi = tile % tiles_x, tile // tiles_x
        c0, r0 = h + txi * TX, h + tyi * TY
        edge = txi == 0 or tyi == 0 or txi == tiles_x - 1 or tyi == tiles_y - 1
        S0 = np.full((S0H, S0W), np.nan); S1 = np.full((S1H, S1W), np.nan)
        for q in range(S0H * S0W):
            a, b = q // S0W, q % S0W
            jj, kk = r0 - 2 + a, c0 - 2 + b
            S0[a][b] = G(sd_old, kk + jj * x) if (0 <= jj < y and 0 <= kk < x) else 0.0
        if pre_halo and (c0 - 2 <= h - 1 or c0 + TX + 1 >= xe or
                         r0 - 2 <= h - 1 or r0 + TY + 1 >= ye):
            reflect_staged_ring(S0, r0, c0, h, xe, ye)
        regs = {}
        for t in range(TX):
            kk = c0 + t
            rk1 = [0.0] * TY; kx0 = [0.0] * TY; kx1 = [0.0] * TY; kyv = [0.0] * (TY + 1)
            for m in range(TY):
                jj = r0 + m
                s1 = 0.0
                if kk < xe and jj < ye:
                    i = kk + jj * x
                    kx0[m] = G(kx, i); kx1[m] = G(kx, i + 1); kyv[m] = G(ky, i); kyv[m + 1] = G(ky, i + x)
                    c = S0[m + 2][t + 2]
                    rn = G(r, i) - op5(c, S0[m + 2][t + 1], S0[m + 2][t + 3], S0[m + 1][t + 2], S0[m + 3][t + 2], kx0[m], kx1[m], kyv[m], kyv[m + 1])
                    rk1[m] = rn
                    s1 = a0 * c + b0 * rn
                S1[m + 1][t + 1] = s1
            regs[t] = (rk1, kx0, kx1, kyv)
        for q in range(2 * S1W + 2 * TY):
            if q < 2 * S1W:
                a = 0 if q < S1W else S1H - 1; b = q if q < S1W else q - S1W
            else:
                e = q - 2 * S1W; a = 1 + (e >> 1); b = S1W - 1 if (e & 1) else 0
            jj, kc = r0 - 1 + a, c0 - 1 + b
            if h <= jj < ye and h <= kc < xe:
                i = kc + jj * x
                c = S0[a + 1][b + 1]
                rn = G(r, i) - op5(c, S0[a + 1][b], S0[a + 1][b + 2], S0[a][b + 1], S0[a + 2][b + 1], G(kx, i), G(kx, i + 1), G(ky, i), G(ky, i + x))
                S1[a][b] = a0 * c + b0 * rn
        if edge:
            upd = {}
            for q in range(S1H * S1W):
                a, b = q // S1W, q % S1W
                jj, kc = r0 - 1 + a, c0 - 1 + b
                ring_box = h - 1 <= jj <= ye and h - 1 <= kc <= xe
                inner = h <= jj < ye and h <= kc < xe
                if not ring_box or inner: continue
                mj = h if jj == h - 1 else (ye - 1 if jj == ye else jj)
                mk = h if kc == h - 1 else (xe - 1 if kc == xe else kc)
                ma, mb = mj - r0 + 1, mk - c0 + 1
                assert 0 <= ma < S1H and 0 <= mb < S1W, ("mirror outside S1", jj, kc)
                mjj, mkk = r0 - 1 + ma, c0 - 1 + mb
                assert h <= mjj < ye and h <= mkk < xe
                upd[(a, b)] = S1[ma][mb]
            for (a, b), v in upd.items(): S1[a][b] = v
        for t in range(TX):
            rk1, kx0, kx1, kyv = regs[t]
            kk = c0 + t
            for m in range(TY):
                jj = r0 + m
                if kk < xe and jj < ye:
                    i = kk + jj * x
                    c = S1[m + 1][t + 1]
                    nb = (S1[m + 1][t], S1[m + 1][t + 2], S1[m][t + 1], S1[m + 2][t + 1])
                    assert not any(np.isnan(v) for v in nb + (c,)), ("unset S1 used", jj, kk)
                    rn = rk1[m] - op5(c, nb[0], nb[1], nb[2], nb[3], kx0[m], kx1[m], kyv[m], kyv[m + 1])
                    P(r_new, i, rn)
                    uo = G(u, i)
                    if p_pend is not None:
                        uo = uo + alpha_pend * G(p_pend, i)
                    P(u, i, (uo + S0[m + 2][t + 2]) + c)
                    P(sd_new, i, a1 * c + b1 * rn); written[i] += 1
        if edge:
            ka = 0 if txi == 0 else c0
            kb = x if txi == tiles_x - 1 else min(c0 + TX, xe)
            ja = 0 if tyi == 0 else r0
            jb = y if tyi == tiles_y - 1 else min(r0 + TY, ye)
            w = kb - ka
            for q in range(w * (jb - ja)):
                jj, kc = ja + q // w, ka + q % w
                if h <= jj < ye and h <= kc < xe: continue
                mj = h if jj == h - 1 else (ye - 1 if jj == ye else jj)
                mk = h if kc == h - 1 else (xe - 1 if kc == xe else kc)
                if h <= mj < ye and h <= mk < xe:
                    a, b = mj - r0 + 1, mk - c0 + 1
                    assert 0 <= a < S1H and 0 <= b < S1W, ("S1 index", jj, kc)
                    v = S1[a][b]
                    assert not np.isnan(v)
                else:
                    v = G(sd_old, mk + mj * x)
                P(sd_new, kc + jj * x, v); written[kc + jj * x] += 1
                P(r_new, kc + jj * x, G(r, kc + jj * x))
    assert (written == 1).all(), ("cells written != once", np.flatnonzero(written != 1)[:10])


def model_cheby(x, y, h, u_old, u_new, u0, p_old, p_new, r, w, kx, ky, a0, b0, a1, b1,
                pre_halo=False):
    n = x * y
    def G(arr, i):
        assert 0 <= i < n, ("read OOB", i); return arr[i]
    def P(arr, i, v):
        assert 0 <= i < n, ("write OOB", i); arr[i] = v
    xe, ye = x - h, y - h
    tiles_x = (x - 2 * h + TX - 1) // TX; tiles_y = (y - 2 * h + TY - 1) // TY
    wu = np.zeros(n, dtype=int)
    for blk in range(tiles_x * tiles_y):
        tile = tile_of_block(blk, tiles_x * tiles_y)
        txi, tyi = tile % tiles_x, tile // tiles_x
        c0, r0 = h + txi * TX, h + tyi * TY
        edge = txi == 0 or tyi == 0 or txi == tiles_x - 1 or tyi == tiles_y - 1
        S0 = np.full((S0H, S0W), np.nan)
        for q in range(S0H * S0W):
            a, b = q // S0W, q % S0W
            jj, kk = r0 - 2 + a, c0 - 2 + b
            S0[a][b] = G(u_old, kk + jj * x) if (0 <= jj < y and 0 <= kk < x) else 0.0
        if pre_halo and (c0 - 2 <= h - 1 or c0 + TX + 1 >= xe or
                         r0 - 2 <= h - 1 or r0 + TY + 1 >= ye):
            reflect_staged_ring(S0, r0, c0, h, xe, ye)
        S1 = np.full((S1H, S1W), np.nan)
        regs = {}
        for t in range(TX):
            kk = c0 + t
            bv = [0.0]*TY; pk = [0.0]*TY; k0 = [0.0]*TY; k1 = [0.0]*TY; kyv = [0.0]*(TY+1)
            for m in range(TY):
                jj = r0 + m
                if kk < xe and jj < ye:
                    i = kk + jj * x
                    bv[m] = G(u0, i); pk[m] = G(p_old, i); k0[m] = G(kx, i); k1[m] = G(kx, i + 1)
                    kyv[m] = G(ky, i); kyv[m + 1] = G(ky, i + x)
                    c = S0[m + 2][t + 2]
                    wv = op5(c, S0[m + 2][t + 1], S0[m + 2][t + 3], S0[m + 1][t + 2], S0[m + 3][t + 2], k0[m], k1[m], kyv[m], kyv[m + 1])
                    rv = bv[m] - wv
                    pk[m] = a0 * pk[m] + b0 * rv
                    S1[m + 1][t + 1] = c + pk[m]
                else:
                    S1[m + 1][t + 1] = 0.0
            regs[t] = (bv, pk, k0, k1, kyv)
        for q in range(2 * S1W + 2 * TY):
            if q < 2 * S1W:
                a = 0 if q < S1W else S1H - 1; b = q if q < S1W else q - S1W
            else:
                e = q - 2 * S1W; a = 1 + (e >> 1); b = S1W - 1 if (e & 1) else 0
            jj, kc = r0 - 1 + a, c0 - 1 + b
            if h <= jj < ye and h <= kc < xe:
                i = kc + jj * x
                c = S0[a + 1][b + 1]
                wv = op5(c, S0[a + 1][b], S0[a + 1][b + 2], S0[a][b + 1], S0[a + 2][b + 1], G(kx, i), G(kx, i + 1), G(ky, i), G(ky, i + x))
                rv = G(u0, i) - wv
                S1[a][b] = c + (a0 * G(p_old, i) + b0 * rv)
        if edge:
            upd = {}
            for q in range(S1H * S1W):
                a, b = q // S1W, q % S1W
                jj, kc = r0 - 1 + a, c0 - 1 + b
                if not (h - 1 <= jj <= ye and h - 1 <= kc <= xe) or (h <= jj < ye and h <= kc < xe): continue
                mj = h if jj == h - 1 else (ye - 1 if jj == ye else jj)
                mk = h if kc == h - 1 else (xe - 1 if kc == xe else kc)
                upd[(a, b)] = S1[mj - r0 + 1][mk - c0 + 1]
            for (a, b), v in upd.items(): S1[a][b] = v
        for t in range(TX):
            bv, pk, k0, k1, kyv = regs[t]
            kk = c0 + t
            for m in range(TY):
                jj = r0 + m
                if kk < xe and jj < ye:
                    i = kk + jj * x
                    c = S1[m + 1][t + 1]
                    nb = (S1[m + 1][t], S1[m + 1][t + 2], S1[m][t + 1], S1[m + 2][t + 1])
                    assert not any(np.isnan(v) for v in nb + (c,))
                    wv = op5(c, nb[0], nb[1], nb[2], nb[3], k0[m], k1[m], kyv[m], kyv[m + 1])
                    rv = bv[m] - wv
                    pv = a1 * pk[m] + b1 * rv
                    P(w, i, wv); P(r, i, rv); P(p_new, i, pv); P(u_new, i, c + pv); wu[i] += 1
        if edge:
            ka = 0 if txi == 0 else c0
            kb = x if txi == tiles_x - 1 else min(c0 + TX, xe)
            ja = 0 if tyi == 0 else r0
            jb = y if tyi == tiles_y - 1 else min(r0 + TY, ye)
            wd = kb - ka
            for q in range(wd * (jb - ja)):
                jj, kc = ja + q // wd, ka + q % wd
                if h <= jj < ye and h <= kc < xe: continue
                mj = h if jj == h - 1 else (ye - 1 if jj == ye else jj)
                mk = h if kc == h - 1 else (xe - 1 if kc == xe else kc)
                if h <= mj < ye and h <= mk < xe:
                    a, b = mj - r0 + 1, mk - c0 + 1
                    assert 0 <= a < S1H and 0 <= b < S1W
                    v = S1[a][b]; assert not np.isnan(v)
                else:
                    v = G(u_old, mk + mj * x)
                P(u_new, kc + jj * x, v); P(p_new, kc + jj * x, G(p_old, kc + jj * x)); wu[kc + jj * x] += 1
    assert (wu == 1).all()


GRIDS = [(3, 2, 2), (10, 7, 2), (TX + 44, 11, 2), (TX, TY, 2), (TX + 1, TY + 1, 2),
         (2 * TX + 1, 2 * TY + 1, 2), (40, 30, 1), (TX + 43, 13, 3)]


@pytest.mark.parametrize("pending,pre", [(False, False), (True, False), (True, True)])
@pytest.mark.parametrize("xc,yc,h", GRIDS)
def test_ppcg_two_step_model(oracle, xc, yc, h, pending, pre):
    """With `pending`, u += alpha p of cg_calc_ur (interior cells) is applied
    by the pass before the steps, as the first pass of a PPCG outer step
    does; with `pre`, the SD update before the first step is held back into
    the pass, which reflects the ring it stages."""
    rng = np.random.default_rng(5)
    x, y = xc + 2 * h, yc + 2 * h
    n = x * y
    f = {k: rng.standard_normal(n) for k in ("sd", "r", "u", "p")}
    f["kx"] = rng.uniform(0, 5, n)
    f["ky"] = rng.uniform(0, 5, n)
    g = {k: v.copy() for k, v in f.items()}
    a0, b0, a1, b1 = 0.3, 0.7, -0.45, 1.2
    alpha = 0.61
    if pending:
        gu = g["u"].reshape(y, x)
        gu[h:y - h, h:x - h] = gu[h:y - h, h:x - h] + alpha * g["p"].reshape(y, x)[h:y - h, h:x - h]
    if pre:
        oracle.halo_field(x, y, h, 1, g["sd"])
    oracle.ppcg_inner_iteration(x, y, h, a0, b0, g)
    oracle.halo_field(x, y, h, 1, g["sd"])
    oracle.ppcg_inner_iteration(x, y, h, a1, b1, g)
    sd_new = np.full(n, np.nan)
    r_new = np.full(n, np.nan)
    u = f["u"].copy()
    model_ppcg(x, y, h, f["sd"], sd_new, f["r"], r_new, u, f["kx"], f["ky"], a0, b0, a1, b1,
               f["p"] if pending else None, alpha, pre)
    for name, got, want in [("sd", sd_new, g["sd"]), ("r", r_new, g["r"]), ("u", u, g["u"])]:
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), name


@pytest.mark.parametrize("pre", [False, True])
@pytest.mark.parametrize("xc,yc,h", GRIDS)
def test_cheby_two_step_model(oracle, xc, yc, h, pre):
    rng = np.random.default_rng(7)
    x, y = xc + 2 * h, yc + 2 * h
    n = x * y
    f = {k: rng.standard_normal(n) for k in ("u", "u0", "p", "r", "w")}
    f["kx"] = rng.uniform(0, 5, n)
    f["ky"] = rng.uniform(0, 5, n)
    g = {k: v.copy() for k, v in f.items()}
    a0, b0, a1, b1 = 0.3, 0.7, -0.45, 1.2
    if pre:
        oracle.halo_field(x, y, h, 1, g["u"])
    oracle.cheby_iterate(x, y, h, a0, b0, g)
    oracle.halo_field(x, y, h, 1, g["u"])
    oracle.cheby_iterate(x, y, h, a1, b1, g)
    u_new = np.full(n, np.nan)
    p_new = np.full(n, np.nan)
    r, w = f["r"].copy(), f["w"].copy()
    model_cheby(x, y, h, f["u"], u_new, f["u0"], f["p"], p_new, r, w, f["kx"], f["ky"],
                a0, b0, a1, b1, pre)
    for name, got, want in [("u", u_new, g["u"]), ("p", p_new, g["p"]), ("r", r, g["r"]),
                            ("w", w, g["w"])]:
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), name
