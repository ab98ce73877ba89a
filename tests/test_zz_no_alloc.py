"""tea_hip_calculate_2norm_noalloc, the entry point that allocates nothing,
against tea_hip_calculate_2norm through the C ABI: bit for bit, on several
calls in a row (a smaller grid after a larger one, so the fixed scratch
array still holds the earlier call's partials), and grids beyond the
scratch rejected.  The result is written to device memory (a chunk field)
on the null stream and read back after a device synchronisation."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libteahip.so itself links (one runtime per process)."""
    from exploringsycl_amd import native
    native.load()
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipDeviceSynchronize.restype = ctypes.c_int
    return hip


def _case(x, y, h, seed):
    from exploringsycl_amd.chunk import DeviceChunk
    rng = np.random.default_rng(seed)
    c = DeviceChunk(x, y, h)
    c.write("u", rng.standard_normal(x * y) * 10.0 ** rng.uniform(-3, 3, x * y))
    c.write("w", np.full(x * y, np.nan))
    return c, c.calculate_2norm("u")


def _noalloc(c):
    """The no-allocation 2-norm of the chunk's u, written into w[0]."""
    from exploringsycl_amd import native
    lib = native.load()
    native.check(lib.tea_hip_flush(c.handle), "flush")
    src, _ = c.field_info("u")
    dst, _ = c.field_info("w")
    native.check(lib.tea_hip_calculate_2norm_noalloc(0, ctypes.c_void_p(src), c.x, c.y,
                                                     c.halo_depth, ctypes.c_void_p(dst),
                                                     None), "calculate_2norm_noalloc")
    assert _hip().hipDeviceSynchronize() == 0
    return float(c.read("w").ravel()[0])


def test_noalloc_2norm_matches_the_chunk_entry_point():
    # Large first, then smaller: 1000 x 700 (171 spans), 300 x 203 (odd x,
    # 15 spans), 10 x 6 (one group of 64), then 8196^2 (the largest grid).
    for (x, y, h), seed in (((1000, 700, 2), 1), ((300, 203, 2), 2), ((10, 6, 2), 3),
                            ((8196, 8196, 2), 4)):
        c, want = _case(x, y, h, seed)
        try:
            got = _noalloc(c)
        finally:
            c.close()
        assert got == want, ((x, y), got, want)


def test_noalloc_2norm_rejects_grids_beyond_the_scratch():
    from exploringsycl_amd import native
    from exploringsycl_amd.chunk import DeviceChunk
    lib = native.load()
    c = DeviceChunk(16, 16, 2)
    try:
        src, _ = c.field_info("u")
        dst, _ = c.field_info("w")
        # 8208^2 cells need more span partials than the scratch holds: the
        # call must fail before launching anything (the small field given
        # is never read).
        rc = lib.tea_hip_calculate_2norm_noalloc(0, ctypes.c_void_p(src), 8208, 8208, 2,
                                                ctypes.c_void_p(dst), None)
        assert rc != 0
        assert b"exceed the scratch" in lib.tea_hip_last_error()
    finally:
        c.close()
