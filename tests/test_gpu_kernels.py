"""Per-kernel parity of libteahip.so against the CPU oracle: every field and
every returned scalar must be bitwise equal (FP64, same operation order,
same reduction tree).  Runs through the C ABI (DeviceChunk)."""

import numpy as np
import pytest

from exploringsycl_amd.native import HALO_SLOTS
from tests.oracle_ctypes import new_fields

pytestmark = pytest.mark.gpu

# (x_cells, y_cells): one group, one span, ragged rows, several spans, and
# enough spans for three final-reduce passes.
SIZES = [(1, 1), (3, 5), (10, 10), (60, 3), (250, 250), (601, 599), (1030, 300)]
BIG = (4096, 4096)
H = 2


def _chunk(x, y):
    from exploringsycl_amd.chunk import DeviceChunk
    return DeviceChunk(x, y, H)


def _random_fields(x, y, seed):
    rng = np.random.default_rng(seed)
    f = new_fields(x, y, H)
    for k in ("density", "energy0", "energy", "u", "u0", "p", "r", "w", "sd",
              "volume"):
        f[k][:] = rng.standard_normal(f[k].size)
    f["density"][:] = rng.uniform(0.1, 100.0, f["density"].size)
    f["volume"][:] = rng.uniform(0.5, 1.5, f["volume"].size)
    f["kx"][:] = rng.uniform(0.0, 50.0, f["kx"].size)
    f["ky"][:] = rng.uniform(0.0, 50.0, f["ky"].size)
    return f


def _upload(c, f, names):
    for n in names:
        c.write(n, f[n])


def _assert_fields(c, f, names):
    for n in names:
        got = c.read(n).ravel()
        want = f[n]
        assert got.shape == want.shape, n
        bad = np.flatnonzero((got.view(np.int64) != want.view(np.int64))
                             & ~(np.isnan(got) & np.isnan(want)))
        assert bad.size == 0, (f"field {n}: {bad.size} cells differ, first at "
                               f"{bad[0]}: {got[bad[0]]!r} vs {want[bad[0]]!r}")


def _same(a, b):
    """Bitwise equal, or both NaN (NaN payloads differ between CPU and GPU)."""
    a, b = np.float64(a), np.float64(b)
    return a.view(np.int64) == b.view(np.int64) or (np.isnan(a) and np.isnan(b))


@pytest.mark.parametrize("cells", SIZES)
def test_set_chunk_data_and_state(oracle, cells):
    xc, yc = cells
    x, y = xc + 2 * H, yc + 2 * H
    dx, dy = 10.0 / xc, 10.0 / yc
    f = new_fields(x, y, H)
    states = [dict(density=100.0, energy=1e-4),
              dict(density=0.1, energy=25.0, x_min=0.0 + dx / 100, x_max=1.0 - dx / 100,
                   y_min=1.0 + dy / 100, y_max=2.0 - dy / 100),
              dict(density=3.0, energy=7.0, x_min=2.5, x_max=7.25, y_min=0.3,
                   y_max=9.1),
              dict(density=0.5, energy=2.0, x_min=5.0, y_min=5.0, x_max=0.0,
                   y_max=0.0, radius=2.2, geometry="circular")]
    with _chunk(x, y) as c:
        c.set_chunk_data(0.0, 0.0, dx, dy)
        c.set_chunk_state(states)
        c.store_energy()
        oracle.set_chunk_data(x, y, H, 0.0, 0.0, dx, dy, f)
        oracle.set_chunk_state(x, y, states, f)
        f["energy"][:] = f["energy0"]
        _assert_fields(c, f, ["vertex_x", "vertex_y", "cell_x", "cell_y", "volume",
                              "x_area", "y_area", "energy0", "density", "u", "energy"])


@pytest.mark.parametrize("cells", SIZES)
@pytest.mark.parametrize("depth", [1, 2])
def test_local_halos(oracle, cells, depth):
    xc, yc = cells
    x, y = xc + 2 * H, yc + 2 * H
    if min(xc, yc) < depth:
        pytest.skip("interior smaller than the halo depth")
    f = _random_fields(x, y, 7)
    names = {"density": "density", "energy0": "energy0", "energy1": "energy",
             "u": "u", "p": "p", "sd": "sd"}
    flagged = ["density", "p", "energy1", "sd"]
    with _chunk(x, y) as c:
        _upload(c, f, names.values())
        c.local_halos(depth, flagged)
        for slot in flagged:
            oracle.halo_field(x, y, H, depth, f[names[slot]])
        _assert_fields(c, f, names.values())


def test_local_halos_partial_faces(oracle):
    x, y = 37 + 4, 29 + 4
    f = _random_fields(x, y, 8)
    for nb in [(-1, 3, -1, 5), (2, -1, 4, -1), (1, 1, -1, -1)]:
        with _chunk(x, y) as c:
            _upload(c, f, ["u"])
            c.local_halos(2, ["u"], neighbours=nb)
            g = f["u"].copy()
            oracle.halo_field(x, y, H, 2, g, neighbours=nb)
            got = c.read("u").ravel()
            assert np.array_equal(got.view(np.int64), g.view(np.int64)), nb


@pytest.mark.parametrize("cells", SIZES)
def test_field_summary(oracle, cells):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    f = _random_fields(x, y, 11)
    with _chunk(x, y) as c:
        _upload(c, f, ["volume", "density", "energy0", "u"])
        got = c.field_summary()
        want = oracle.field_summary(x, y, H, f)
        assert all(_same(a, b) for a, b in zip(got, want)), (got, want)


@pytest.mark.parametrize("cells", SIZES)
@pytest.mark.parametrize("coef", [1, 2])
def test_cg_init(oracle, cells, coef):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    f = _random_fields(x, y, 13)
    names = ["p", "r", "u", "w", "kx", "ky", "density", "energy"]
    with _chunk(x, y) as c:
        _upload(c, f, names)
        got = c.cg_init(coef, 0.7, 1.3)
        want = oracle.cg_init(x, y, H, coef, 0.7, 1.3, f)
        assert _same(got, want), (got, want)
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", SIZES + [BIG])
def test_cg_calc_w_ur_p_unfused(oracle, cells):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    f = _random_fields(x, y, 17)
    names = ["p", "r", "u", "w", "kx", "ky"]
    with _chunk(x, y) as c:
        c.set_fused(False)
        _upload(c, f, names)
        pw = c.cg_calc_w()
        assert _same(pw, oracle.cg_calc_w(x, y, H, f))
        rrn = c.cg_calc_ur(0.37)
        assert _same(rrn, oracle.cg_calc_ur(x, y, H, 0.37, f))
        c.cg_calc_p(0.81)
        oracle.cg_calc_p(x, y, H, 0.81, f)
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", SIZES)
def test_shared_solver_kernels(oracle, cells):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    f = _random_fields(x, y, 19)
    names = ["u", "u0", "r", "kx", "ky", "density", "energy"]
    with _chunk(x, y) as c:
        _upload(c, f, names)
        c.calculate_residual()
        oracle.calculate_residual(x, y, H, f)
        assert _same(c.calculate_2norm("r"), oracle.calculate_2norm(x, y, H, f["r"]))
        c.copy_u()
        oracle.copy_u(x, y, H, f)
        c.finalise()
        oracle.finalise(x, y, H, f)
        _assert_fields(c, f, names)


def _cg_sequence(c, oracle, x, y, f, iters, coef=2):
    """The driver's call sequence (TeaLeaf/drivers/cg_driver.c:31-124) on the
    device and on the oracle; scalars must agree bitwise at every step."""
    rro = c.cg_init(coef, 0.9, 1.1)
    rro_o = oracle.cg_init(x, y, H, coef, 0.9, 1.1, f)
    assert _same(rro, rro_o)
    c.local_halos(1, ["u", "p"])
    for n in ("u", "p"):
        oracle.halo_field(x, y, H, 1, f[n])
    c.copy_u()
    oracle.copy_u(x, y, H, f)
    for it in range(iters):
        pw = c.cg_calc_w()
        pw_o = oracle.cg_calc_w(x, y, H, f)
        assert _same(pw, pw_o), (it, pw, pw_o)
        with np.errstate(all="ignore"):  # tiny grids converge to rr == 0
            alpha = float(np.float64(rro) / np.float64(pw))
        rrn = c.cg_calc_ur(alpha)
        rrn_o = oracle.cg_calc_ur(x, y, H, alpha, f)
        assert _same(rrn, rrn_o), (it, rrn, rrn_o)
        with np.errstate(all="ignore"):
            beta = float(np.float64(rrn) / np.float64(rro))
        c.cg_calc_p(beta)
        oracle.cg_calc_p(x, y, H, beta, f)
        c.local_halos(1, ["u", "p"])
        for n in ("u", "p"):
            oracle.halo_field(x, y, H, 1, f[n])
        rro = rrn
    return rro


@pytest.mark.parametrize("cells", [(1, 1), (10, 10), (60, 3), (250, 250), (601, 599)])
@pytest.mark.parametrize("fused", [True, False])
def test_cg_iterations_match_oracle(oracle, cells, fused):
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    f = _random_fields(x, y, 23)
    f["density"][:] = np.random.default_rng(5).uniform(0.1, 100.0, x * y)
    f["energy"][:] = np.random.default_rng(6).uniform(0.1, 25.0, x * y)
    for n in ("density", "energy"):
        oracle.halo_field(x, y, H, 2, f[n])
    names = ["p", "r", "u", "u0", "w", "kx", "ky", "density", "energy"]
    with _chunk(x, y) as c:
        c.set_fused(fused)
        _upload(c, f, names)
        _cg_sequence(c, oracle, x, y, f, iters=25)
        # Reading applies any deferred work; every cell must match.
        _assert_fields(c, f, names)
        c.calculate_residual()
        oracle.calculate_residual(x, y, H, f)
        _assert_fields(c, f, ["r"])


def test_fused_large_grid_matches_oracle(oracle):
    xc = yc = 1024
    x, y = xc + 2 * H, yc + 2 * H
    f = _random_fields(x, y, 29)
    f["density"][:] = np.random.default_rng(1).uniform(0.1, 100.0, x * y)
    f["energy"][:] = np.random.default_rng(2).uniform(0.1, 25.0, x * y)
    names = ["p", "r", "u", "u0", "w", "kx", "ky", "density", "energy"]
    with _chunk(x, y) as c:
        _upload(c, f, names)
        _cg_sequence(c, oracle, x, y, f, iters=10)
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", [(250, 250), (601, 599)])
def test_cg_iterations_from_unreflected_fields(oracle, cells):
    """CG iterations on random fields whose halo cells were never reflected.
    cg_calc_p's U/P halo update is folded into the next fused cg_calc_w only
    once a real update has made the face lines beyond the ring (copied from
    cells no iteration changes) what a reflective update leaves: here the
    first update runs for real and the later ones are folded."""
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    f = _random_fields(x, y, 29)
    names = ["p", "w", "u", "r", "kx", "ky"]
    with _chunk(x, y) as c:
        _upload(c, f, names)
        for alpha, beta in [(0.6, 0.3), (0.45, 0.8), (0.2, 0.5)]:
            assert _same(c.cg_calc_w(), oracle.cg_calc_w(x, y, H, f))
            assert _same(c.cg_calc_ur(alpha), oracle.cg_calc_ur(x, y, H, alpha, f))
            c.cg_calc_p(beta)
            oracle.cg_calc_p(x, y, H, beta, f)
            c.local_halos(1, ["u", "p"])
            oracle.halo_field(x, y, H, 1, f["u"])
            oracle.halo_field(x, y, H, 1, f["p"])
        assert _same(c.cg_calc_w(), oracle.cg_calc_w(x, y, H, f))
        _assert_fields(c, f, names)


@pytest.mark.parametrize("cells", [(250, 250), (4096, 16)])
def test_reductions_finish_with_the_wait_sentinel_as_input(oracle, cells):
    """The host waits for a reduction's partials by polling for a sentinel
    (a signalling NaN) to be overwritten.  A field holding exactly that bit
    pattern still gives a (quiet) NaN norm, returned without hanging, and the
    next reduction is exact again."""
    x, y = cells[0] + 2 * H, cells[1] + 2 * H
    f = _random_fields(x, y, 23)
    snan = np.array([0x7FF4C0FFEE5E4D1D], dtype=np.uint64).view(np.float64)[0]
    bad = f["r"].copy()
    bad[(y // 2) * x + x // 2] = snan
    with _chunk(x, y) as c:
        c.write("r", bad)
        got = c.calculate_2norm("r")
        assert np.isnan(got)
        assert np.array(got).view(np.uint64) != np.uint64(0x7FF4C0FFEE5E4D1D)
        c.write("r", f["r"])
        assert _same(c.calculate_2norm("r"), oracle.calculate_2norm(x, y, H, f["r"]))


def test_two_norm_rejects_foreign_pointer():
    from exploringsycl_amd.native import TeaHipError
    import ctypes
    with _chunk(14, 14) as c:
        v = ctypes.c_double()
        st = c.lib.tea_hip_calculate_2norm(c.handle, ctypes.c_void_p(16), ctypes.byref(v))
        assert st != 0
        assert b"not an x*y field" in c.lib.tea_hip_last_error()
        with pytest.raises(TeaHipError):
            c.local_halos(3, ["u"])  # deeper than the halo


def test_halo_slot_names():
    assert HALO_SLOTS == ["density", "energy0", "energy1", "u", "p", "sd"]


@pytest.mark.parametrize("xc", [123, 124, 126])
def test_mixed_precision_kernels_match_oracle(oracle, xc):
    """The mixed-precision CG entry points against their CPU restatement,
    call by call, on ragged grids (bitwise FP32 arrays and FP64 scalars):
    x = 127 takes the one-cell-per-lane kernels, x = 128 the four-cell
    kernels (21 pair blocks: the last four-cell block is half empty), x = 130
    the two-cell kernels."""
    import ctypes
    yc = 77
    x, y = xc + 2 * H, yc + 2 * H
    f = _random_fields(x, y, 31)
    f["density"][:] = np.random.default_rng(3).uniform(0.1, 100.0, x * y)
    f["energy"][:] = np.random.default_rng(4).uniform(0.1, 25.0, x * y)
    for n in ("density", "energy"):
        oracle.halo_field(x, y, H, 2, f[n])
    names = ["p", "r", "u", "u0", "w", "kx", "ky", "density", "energy"]
    L = oracle.lib
    F = ctypes.POINTER(ctypes.c_float)
    P = ctypes.POINTER(ctypes.c_double)
    for fn, args in {
            "orc_mpcg_convert": [ctypes.c_int] * 3 + [P] * 5 + [F] * 4 + [P],
            "orc_mpcg_calc_w": [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int,
                                                     ctypes.c_double, ctypes.c_int]
                               + [F, F, F, P, F, F, F],
            "orc_mpcg_calc_r": [ctypes.c_int] * 3 + [ctypes.c_double, F, F],
            "orc_mpcg_refresh": [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int,
                                                      P, F, P, P, P, F, F, P],
            "orc_dot32": [ctypes.c_int] * 3 + [F, F]}.items():
        getattr(L, fn).argtypes = args
    for fn in ("orc_mpcg_calc_w", "orc_mpcg_calc_r", "orc_dot32"):
        getattr(L, fn).restype = ctypes.c_double
    n = x * y
    mp = {k: np.zeros(n, dtype=np.float32) for k in ("p0", "p1", "r", "w", "kx", "ky")}

    def fp(a):
        return a.ctypes.data_as(F)

    def dp(a):
        return a.ctypes.data_as(P)

    with _chunk(x, y) as c:
        _upload(c, f, names)
        rr, bb = c.mpcg_init(2, 0.9, 1.1)
        oracle.cg_init(x, y, H, 2, 0.9, 1.1, f)
        for nm in ("u", "p"):
            oracle.halo_field(x, y, H, 1, f[nm])
        oracle.copy_u(x, y, H, f)
        out = np.zeros(3)
        L.orc_mpcg_convert(x, y, H, dp(f["kx"]), dp(f["ky"]), dp(f["p"]), dp(f["r"]),
                           dp(f["u0"]), fp(mp["kx"]), fp(mp["ky"]), fp(mp["p0"]),
                           fp(mp["r"]), dp(out))
        mp["p1"][:] = mp["p0"]
        assert _same(rr, out[0]) and _same(bb, out[1])
        cur = 0
        alpha = beta = 0.0
        pu = pp = False
        for it in range(30):
            # The flexible recursion of the mixed-precision driver:
            # alpha = p.r / p.w, beta = -r.w / p.w.
            pw, pr = c.mpcg_calc_w(alpha, pu, beta, pp)
            po, pn = (mp["p0"], mp["p1"]) if cur == 0 else (mp["p1"], mp["p0"])
            pw_o = L.orc_mpcg_calc_w(x, y, H, alpha, int(pu), beta, int(pp), fp(po),
                                     fp(pn), fp(mp["r"]), dp(f["u"]), fp(mp["w"]),
                                     fp(mp["kx"]), fp(mp["ky"]))
            if pp:
                cur ^= 1
            assert _same(pw, pw_o), (it, pw, pw_o)
            pcur = mp["p0"] if cur == 0 else mp["p1"]
            assert _same(pr, L.orc_dot32(x, y, H, fp(pcur), fp(mp["r"]))), it
            alpha = pr / pw
            rrn, rw = c.mpcg_calc_r(alpha)
            assert _same(rrn, L.orc_mpcg_calc_r(x, y, H, alpha, fp(mp["r"]), fp(mp["w"])))
            assert _same(rw, L.orc_dot32(x, y, H, fp(mp["r"]), fp(mp["w"]))), it
            pu = True
            if it % 7 == 6:
                a, b, rw = c.mpcg_refresh(alpha, True)
                L.orc_mpcg_refresh(x, y, H, alpha, 1, dp(f["u"]), fp(pcur), dp(f["u0"]),
                                   dp(f["kx"]), dp(f["ky"]), fp(mp["r"]), fp(mp["w"]),
                                   dp(out))
                assert _same(a, out[0]) and _same(b, out[1]) and _same(rw, out[2])
                pu = False
            beta = -rw / pw
            pp = True
        for nm, arr in (("r", mp["r"]), ("w", mp["w"]),
                        ("p", mp["p0"] if cur == 0 else mp["p1"])):
            got = c.read_mp(nm).ravel()
            assert np.array_equal(got.view(np.int32), arr.view(np.int32)), nm
        _assert_fields(c, f, ["u"])


def test_copy_bandwidth_is_a_plausible_rate():
    """tea_hip_copy_bandwidth: a 256 MiB 16-byte device copy reports a rate
    between 1 TB/s and the 8 TB/s peak; bad arguments are refused."""
    import ctypes
    from exploringsycl_amd import native
    lib = native.load()
    gbs = ctypes.c_double()
    assert lib.tea_hip_copy_bandwidth(0, 256 << 20, 5, ctypes.byref(gbs)) == 0
    assert 1000.0 < gbs.value < 8000.0, gbs.value
    assert lib.tea_hip_copy_bandwidth(0, 17, 1, ctypes.byref(gbs)) != 0
    assert lib.tea_hip_copy_bandwidth(0, 256, 1, None) != 0


@pytest.mark.parametrize("cells", SIZES)
@pytest.mark.parametrize("depth", [1, 2])
def test_pack_or_unpack(oracle, cells, depth):
    """Remote-halo strips (kernel_interface.cpp:129-170): every face, both
    directions, through the host buffer."""
    xc, yc = cells
    x, y = xc + 2 * H, yc + 2 * H
    f = _random_fields(x, y, 31)
    with _chunk(x, y) as c:
        _upload(c, f, ["p"])
        for face in range(4):
            n = (y if face < 2 else x) * depth
            got = np.zeros(n)
            want = np.zeros(n)
            c.pack_or_unpack(depth, face, True, "p", got)
            oracle.pack_or_unpack(x, y, H, depth, face, True, f["p"], want)
            assert np.array_equal(got.view(np.int64), want.view(np.int64)), face
            strip = np.random.default_rng(face).standard_normal(n)
            c.pack_or_unpack(depth, face, False, "p", strip)
            oracle.pack_or_unpack(x, y, H, depth, face, False, f["p"], strip.copy())
        _assert_fields(c, f, ["p"])
