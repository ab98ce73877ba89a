"""One TeaLeaf chunk on the GPU, driven through libteahip.so.

Method names follow the reference's kernel_interface.h entry points
(TeaLeaf/kernel_interface.h:13-71) without the ``run_`` prefix; scalar
results are returned instead of accumulated into out-parameters.
"""

import ctypes

import numpy as np

from . import native
from .native import FIELD_ID, HALO_SLOTS, check


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class DeviceChunk:
    """Device state of one chunk of ``x`` by ``y`` cells, halo included."""

    def __init__(self, x, y, halo_depth=2, device=0, fused=True):
        self.lib = native.load()
        self.x, self.y, self.halo_depth = x, y, halo_depth
        h = ctypes.c_void_p()
        check(self.lib.tea_hip_chunk_create(device, x, y, halo_depth,
                                            ctypes.byref(h)), "chunk_create")
        self.handle = h
        if not fused:
            self.set_fused(False)

    def close(self):
        if self.handle:
            check(self.lib.tea_hip_chunk_destroy(self.handle), "chunk_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields ------------------------------------------------------------
    def field_info(self, name):
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        check(self.lib.tea_hip_field(self.handle, FIELD_ID[name],
                                     ctypes.byref(p), ctypes.byref(n)), "field")
        return p.value, n.value

    def read(self, name):
        _, n = self.field_info(name)
        out = np.empty(n, dtype=np.float64)
        check(self.lib.tea_hip_field_read(self.handle, FIELD_ID[name],
                                          _ptr(out), n), f"read {name}")
        if n == self.x * self.y:
            return out.reshape(self.y, self.x)
        return out

    def write(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        _, n = self.field_info(name)
        if a.size > n:
            raise ValueError(f"{name}: {a.size} values for a field of {n}")
        check(self.lib.tea_hip_field_write(self.handle, FIELD_ID[name],
                                           _ptr(a), a.size), f"write {name}")

    # ---- kernel_interface entry points ---------------------------------------
    def set_chunk_data(self, x_min, y_min, dx, dy):
        check(self.lib.tea_hip_set_chunk_data(self.handle, x_min, y_min, dx, dy),
              "set_chunk_data")

    def set_chunk_state(self, states):
        """``states``: list of dicts with density, energy and (for states
        after the first) x_min, y_min, x_max, y_max, radius, geometry."""
        n = len(states)
        cols = {k: np.array([float(s.get(k, 0.0)) for s in states])
                for k in ("density", "energy", "x_min", "y_min", "x_max",
                          "y_max", "radius")}
        geom = np.array([native.GEOMETRY[s.get("geometry", "rectangle")]
                         for s in states], dtype=np.int32)
        check(self.lib.tea_hip_set_chunk_state(
            self.handle, n, *[_ptr(cols[k]) for k in
                              ("density", "energy", "x_min", "y_min", "x_max",
                               "y_max", "radius")],
            geom.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
            "set_chunk_state")

    def store_energy(self):
        check(self.lib.tea_hip_store_energy(self.handle), "store_energy")

    def local_halos(self, depth, fields, neighbours=(-1, -1, -1, -1)):
        nb = (ctypes.c_int * 4)(*neighbours)
        fl = (ctypes.c_int * 6)(*[1 if s in fields else 0 for s in HALO_SLOTS])
        check(self.lib.tea_hip_local_halos(self.handle, depth, nb, fl),
              "local_halos")

    def field_summary(self):
        v = [ctypes.c_double() for _ in range(4)]
        check(self.lib.tea_hip_field_summary(self.handle,
                                             *[ctypes.byref(a) for a in v]),
              "field_summary")
        return tuple(a.value for a in v)

    def cg_init(self, coefficient, rx, ry):
        v = ctypes.c_double()
        check(self.lib.tea_hip_cg_init(self.handle, coefficient, rx, ry,
                                       ctypes.byref(v)), "cg_init")
        return v.value

    def cg_calc_w(self):
        v = ctypes.c_double()
        check(self.lib.tea_hip_cg_calc_w(self.handle, ctypes.byref(v)), "cg_calc_w")
        return v.value

    def cg_calc_ur(self, alpha):
        v = ctypes.c_double()
        check(self.lib.tea_hip_cg_calc_ur(self.handle, alpha, ctypes.byref(v)),
              "cg_calc_ur")
        return v.value

    def cg_calc_p(self, beta):
        check(self.lib.tea_hip_cg_calc_p(self.handle, beta), "cg_calc_p")

    def copy_u(self):
        check(self.lib.tea_hip_copy_u(self.handle), "copy_u")

    def calculate_residual(self):
        check(self.lib.tea_hip_calculate_residual(self.handle), "calculate_residual")

    def calculate_2norm(self, name):
        p, _ = self.field_info(name)
        v = ctypes.c_double()
        check(self.lib.tea_hip_calculate_2norm(self.handle, ctypes.c_void_p(p),
                                               ctypes.byref(v)), "calculate_2norm")
        return v.value

    def finalise(self):
        check(self.lib.tea_hip_finalise(self.handle), "finalise")

    def pack_or_unpack(self, depth, face, pack, name, buffer):
        """Remote-halo strip of field `name` to (pack) or from the float64
        array `buffer` (kernel_interface.cpp:129-170)."""
        p, _ = self.field_info(name)
        assert buffer.dtype == np.float64 and buffer.flags["C_CONTIGUOUS"]
        check(self.lib.tea_hip_pack_or_unpack(self.handle, depth, face, 1 if pack else 0,
                                              ctypes.c_void_p(p), _ptr(buffer)),
              "pack_or_unpack")

    # ---- Jacobi, Chebyshev, PPCG (kernel_interface.cpp:246-328) ---------------
    def jacobi_init(self, coefficient, rx, ry):
        check(self.lib.tea_hip_jacobi_init(self.handle, coefficient, rx, ry),
              "jacobi_init")

    def jacobi_iterate(self):
        v = ctypes.c_double()
        check(self.lib.tea_hip_jacobi_iterate(self.handle, ctypes.byref(v)),
              "jacobi_iterate")
        return v.value

    def cheby_init(self, theta):
        check(self.lib.tea_hip_cheby_init(self.handle, theta), "cheby_init")

    def cheby_iterate(self, alpha, beta):
        check(self.lib.tea_hip_cheby_iterate(self.handle, alpha, beta),
              "cheby_iterate")

    def ppcg_init(self, theta):
        check(self.lib.tea_hip_ppcg_init(self.handle, theta), "ppcg_init")

    def ppcg_inner_iteration(self, alpha, beta):
        check(self.lib.tea_hip_ppcg_inner_iteration(self.handle, alpha, beta),
              "ppcg_inner_iteration")

    # ---- mixed-precision CG (no reference counterpart) -------------------------
    def mpcg_init(self, coefficient, rx, ry):
        rr, bb = ctypes.c_double(), ctypes.c_double()
        check(self.lib.tea_hip_mpcg_init(self.handle, coefficient, rx, ry,
                                         ctypes.byref(rr), ctypes.byref(bb)),
              "mpcg_init")
        return rr.value, bb.value

    def mpcg_calc_w(self, alpha, apply_u, beta, apply_p):
        """Returns (p.w, p.r)."""
        pw, pr = ctypes.c_double(), ctypes.c_double()
        check(self.lib.tea_hip_mpcg_calc_w(self.handle, alpha, int(apply_u), beta,
                                           int(apply_p), ctypes.byref(pw),
                                           ctypes.byref(pr)),
              "mpcg_calc_w")
        return pw.value, pr.value

    def mpcg_calc_r(self, alpha):
        """Returns (r.r, r.w) of the updated r."""
        rr, rw = ctypes.c_double(), ctypes.c_double()
        check(self.lib.tea_hip_mpcg_calc_r(self.handle, alpha, ctypes.byref(rr),
                                           ctypes.byref(rw)),
              "mpcg_calc_r")
        return rr.value, rw.value

    def mpcg_refresh(self, alpha, apply_u):
        """Returns (r.r of the FP32 residual, r.r of the FP64 one, r.w)."""
        a, b, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.tea_hip_mpcg_refresh(self.handle, alpha, int(apply_u),
                                            ctypes.byref(a), ctypes.byref(b),
                                            ctypes.byref(w)),
              "mpcg_refresh")
        return a.value, b.value, w.value

    def read_mp(self, name):
        out = np.empty(self.x * self.y, dtype=np.float32)
        check(self.lib.tea_hip_mp_read(self.handle, native.MP_FIELDS[name],
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       out.size), f"read_mp {name}")
        return out.reshape(self.y, self.x)

    # ---- flexible CG (no reference counterpart) --------------------------------
    def fcg_attach(self, params=None, norm=None, coarsen=0, levels=2):
        """Allocates the FCG state; with ``params`` also loads the network on
        up to ``levels`` coarse levels."""
        fp = ctypes.POINTER(ctypes.c_float)
        if params is None:
            check(self.lib.tea_hip_fcg_attach(self.handle, None, 0, None, 0, 0), "fcg_attach")
            return
        self._fcg_params = np.ascontiguousarray(params, dtype=np.float32)
        self._fcg_norm = np.ascontiguousarray(norm, dtype=np.float32)
        check(self.lib.tea_hip_fcg_attach(self.handle, self._fcg_params.ctypes.data_as(fp),
                                          self._fcg_params.size,
                                          self._fcg_norm.ctypes.data_as(fp), int(coarsen),
                                          int(levels)),
              "fcg_attach")

    def fcg_levels(self):
        v = ctypes.c_int()
        check(self.lib.tea_hip_fcg_levels(self.handle, ctypes.byref(v)), "fcg_levels")
        return v.value

    def fcg_init(self, coefficient, rx, ry):
        rr, bb = ctypes.c_double(), ctypes.c_double()
        check(self.lib.tea_hip_fcg_init(self.handle, coefficient, rx, ry,
                                        ctypes.byref(rr), ctypes.byref(bb)), "fcg_init")
        return rr.value, bb.value

    def fcg_precondition(self, use_net):
        v = ctypes.c_double()
        check(self.lib.tea_hip_fcg_precondition(self.handle, int(use_net), ctypes.byref(v)),
              "fcg_precondition")
        return v.value

    def fcg_calc_r_precondition(self, alpha, use_net):
        """r -= alpha w, then z = M(r); returns (r.r, z.w)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self.lib.tea_hip_fcg_calc_r_precondition(self.handle, alpha, int(use_net),
                                                       ctypes.byref(a), ctypes.byref(b)),
              "fcg_calc_r_precondition")
        return a.value, b.value

    def fcg_calc_w(self, alpha, apply_u, beta):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self.lib.tea_hip_fcg_calc_w(self.handle, alpha, int(apply_u), beta,
                                          ctypes.byref(a), ctypes.byref(b)), "fcg_calc_w")
        return a.value, b.value

    def read_z(self):
        out = np.empty(self.x * self.y, dtype=np.float32)
        check(self.lib.tea_hip_fcg_read_z(self.handle,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                          out.size), "read_z")
        return out.reshape(self.y, self.x)

    # ---- mode and timing -------------------------------------------------------
    def set_fused(self, on):
        check(self.lib.tea_hip_set_fused(self.handle, 1 if on else 0), "set_fused")

    def flush(self):
        check(self.lib.tea_hip_flush(self.handle), "flush")

    def timing(self, on=True):
        check(self.lib.tea_hip_timing_enable(self.handle, 1 if on else 0), "timing")

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(native.kernel_names()):
            n = ctypes.c_long()
            ms = ctypes.c_double()
            check(self.lib.tea_hip_kernel_stats(self.handle, k, ctypes.byref(n),
                                                ctypes.byref(ms)), "kernel_stats")
            if n.value:
                out[name] = (n.value, ms.value)
        return out
