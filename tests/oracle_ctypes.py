"""ctypes binding of the CPU oracle (oracle/build/libtea_oracle.so) for the
tests.  Test infrastructure only: the oracle is the checker."""

import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(REPO, "oracle", "build", "libtea_oracle.so")
ORACLE_TEALEAF = os.path.join(REPO, "oracle", "build", "tealeaf_oracle")
REFHOST_ORACLE = os.path.join(REPO, "oracle", "_ref", "tealeaf_refhost_oracle")
REFHOST_HIP = os.path.join(REPO, "oracle", "_ref", "tealeaf_refhost_hip")

_D = ctypes.c_double
_P = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


class Oracle:
    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built: make -C oracle")
        lib = ctypes.CDLL(path)
        sig = {
            "orc_tree_sum": ([_P, ctypes.c_size_t], _D),
            "orc_tree64": ([_P], _D),
            "orc_set_chunk_data": ([_I, _I, _I, _D, _D, _D, _D] + [_P] * 7, None),
            "orc_set_chunk_initial_state": ([_I, _I, _D, _D, _P, _P], None),
            "orc_set_chunk_state": ([_I, _I, _I] + [_D] * 7 + [_P] * 7, None),
            "orc_store_energy": ([_I, _I, _P, _P], None),
            "orc_halo_field": ([_I, _I, _I, _I, ctypes.POINTER(_I), _P], None),
            "orc_field_summary": ([_I, _I, _I, _P, _P, _P, _P, _P], None),
            "orc_cg_init": ([_I, _I, _I, _I, _D, _D] + [_P] * 8, _D),
            "orc_cg_calc_w": ([_I, _I, _I, _P, _P, _P, _P], _D),
            "orc_cg_calc_ur": ([_I, _I, _I, _D, _P, _P, _P, _P], _D),
            "orc_cg_calc_p": ([_I, _I, _I, _D, _P, _P], None),
            "orc_copy_u": ([_I, _I, _I, _P, _P], None),
            "orc_calculate_residual": ([_I, _I, _I] + [_P] * 5, None),
            "orc_calculate_2norm": ([_I, _I, _I, _P], _D),
            "orc_finalise": ([_I, _I, _I, _P, _P, _P], None),
            "orc_smvp": ([_P, _P, _P, ctypes.c_long, ctypes.c_long], _D),
            "orc_pack_or_unpack": ([_I] * 6 + [_P, _P], None),
            "orc_jacobi_init": ([_I, _I, _I, _I, _D, _D] + [_P] * 6, None),
            "orc_jacobi_iterate": ([_I, _I, _I] + [_P] * 5, _D),
            "orc_cheby_init": ([_I, _I, _I, _D] + [_P] * 7, None),
            "orc_cheby_iterate": ([_I, _I, _I, _D, _D] + [_P] * 7, None),
            "orc_ppcg_init": ([_I, _I, _I, _D, _P, _P], None),
            "orc_ppcg_inner_iteration": ([_I, _I, _I, _D, _D] + [_P] * 5, None),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        self.lib = lib

    # Thin wrappers taking numpy arrays (modified in place like the kernels).
    def tree_sum(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        return self.lib.orc_tree_sum(_p(v), v.size)

    def set_chunk_data(self, x, y, h, x_min, y_min, dx, dy, f):
        self.lib.orc_set_chunk_data(x, y, h, x_min, y_min, dx, dy, _p(f["vertex_x"]),
                                    _p(f["vertex_y"]), _p(f["cell_x"]), _p(f["cell_y"]),
                                    _p(f["volume"]), _p(f["x_area"]), _p(f["y_area"]))

    def set_chunk_state(self, x, y, states, f):
        s0 = states[0]
        self.lib.orc_set_chunk_initial_state(x, y, s0["energy"], s0["density"],
                                             _p(f["energy0"]), _p(f["density"]))
        geo = {"rectangle": 0, "circular": 1, "point": 2}
        for s in states[1:]:
            self.lib.orc_set_chunk_state(
                x, y, geo[s.get("geometry", "rectangle")], s["x_min"], s["y_min"],
                s["x_max"], s["y_max"], s.get("radius", 0.0), s["density"], s["energy"],
                _p(f["energy0"]), _p(f["density"]), _p(f["u"]), _p(f["cell_x"]),
                _p(f["cell_y"]), _p(f["vertex_x"]), _p(f["vertex_y"]))

    def halo_field(self, x, y, h, depth, b, neighbours=(-1, -1, -1, -1)):
        nb = (_I * 4)(*neighbours)
        self.lib.orc_halo_field(x, y, h, depth, nb, _p(b))

    def field_summary(self, x, y, h, f):
        out = np.zeros(4)
        self.lib.orc_field_summary(x, y, h, _p(f["volume"]), _p(f["density"]),
                                   _p(f["energy0"]), _p(f["u"]), _p(out))
        return tuple(out)

    def cg_init(self, x, y, h, coef, rx, ry, f):
        return self.lib.orc_cg_init(x, y, h, coef, rx, ry, _p(f["p"]), _p(f["r"]),
                                    _p(f["u"]), _p(f["w"]), _p(f["kx"]), _p(f["ky"]),
                                    _p(f["density"]), _p(f["energy"]))

    def cg_calc_w(self, x, y, h, f):
        return self.lib.orc_cg_calc_w(x, y, h, _p(f["w"]), _p(f["p"]), _p(f["kx"]),
                                      _p(f["ky"]))

    def cg_calc_ur(self, x, y, h, alpha, f):
        return self.lib.orc_cg_calc_ur(x, y, h, alpha, _p(f["u"]), _p(f["r"]),
                                       _p(f["p"]), _p(f["w"]))

    def cg_calc_p(self, x, y, h, beta, f):
        self.lib.orc_cg_calc_p(x, y, h, beta, _p(f["p"]), _p(f["r"]))

    def copy_u(self, x, y, h, f):
        self.lib.orc_copy_u(x, y, h, _p(f["u"]), _p(f["u0"]))

    def calculate_residual(self, x, y, h, f):
        self.lib.orc_calculate_residual(x, y, h, _p(f["u"]), _p(f["u0"]), _p(f["r"]),
                                        _p(f["kx"]), _p(f["ky"]))

    def calculate_2norm(self, x, y, h, b):
        return self.lib.orc_calculate_2norm(x, y, h, _p(b))

    def finalise(self, x, y, h, f):
        self.lib.orc_finalise(x, y, h, _p(f["u"]), _p(f["density"]), _p(f["energy"]))

    def pack_or_unpack(self, x, y, h, depth, face, pack, field, buffer):
        self.lib.orc_pack_or_unpack(x, y, h, depth, face, 1 if pack else 0, _p(field),
                                    _p(buffer))

    def jacobi_init(self, x, y, h, coef, rx, ry, f):
        self.lib.orc_jacobi_init(x, y, h, coef, rx, ry, _p(f["u"]), _p(f["u0"]),
                                 _p(f["density"]), _p(f["energy"]), _p(f["kx"]),
                                 _p(f["ky"]))

    def jacobi_iterate(self, x, y, h, f):
        return self.lib.orc_jacobi_iterate(x, y, h, _p(f["u"]), _p(f["u0"]), _p(f["r"]),
                                           _p(f["kx"]), _p(f["ky"]))

    def cheby_init(self, x, y, h, theta, f):
        self.lib.orc_cheby_init(x, y, h, theta, _p(f["p"]), _p(f["r"]), _p(f["u"]),
                                _p(f["u0"]), _p(f["w"]), _p(f["kx"]), _p(f["ky"]))

    def cheby_iterate(self, x, y, h, alpha, beta, f):
        self.lib.orc_cheby_iterate(x, y, h, alpha, beta, _p(f["p"]), _p(f["r"]),
                                   _p(f["u"]), _p(f["u0"]), _p(f["w"]), _p(f["kx"]),
                                   _p(f["ky"]))

    def ppcg_init(self, x, y, h, theta, f):
        self.lib.orc_ppcg_init(x, y, h, theta, _p(f["sd"]), _p(f["r"]))

    def ppcg_inner_iteration(self, x, y, h, alpha, beta, f):
        self.lib.orc_ppcg_inner_iteration(x, y, h, alpha, beta, _p(f["sd"]), _p(f["r"]),
                                          _p(f["u"]), _p(f["kx"]), _p(f["ky"]))


def new_fields(x, y, h=2):
    """Zeroed host fields with the backend's lengths."""
    xy = x * y
    lens = dict(density0=xy, density=xy, energy0=xy, energy=xy, u=xy, u0=xy, p=xy,
                r=xy, mi=xy, w=xy, kx=xy, ky=xy, sd=xy, volume=xy,
                x_area=(x + 1) * y, y_area=x * (y + 1), cell_x=x, cell_y=y,
                cell_dx=x, cell_dy=y, vertex_dx=x + 1, vertex_dy=y + 1,
                vertex_x=x + 1, vertex_y=y + 1, comms_buffer=max(x, y) * h)
    return {k: np.zeros(n) for k, n in lens.items()}


class OracleSolver:
    """The CG path of TeaLeaf/drivers/cg_driver.c on the oracle, for one
    chunk with all faces external (tests only)."""

    def __init__(self, oracle, x_cells, y_cells, states, coefficient=1, h=2,
                 dt=0.004, extent=10.0):
        self.o = oracle
        self.x, self.y, self.h = x_cells + 2 * h, y_cells + 2 * h, h
        self.dx = extent / x_cells
        self.dy = extent / y_cells
        self.rx = dt / (self.dx * self.dx)
        self.ry = dt / (self.dy * self.dy)
        self.coef = coefficient
        self.f = new_fields(self.x, self.y, h)
        oracle.set_chunk_data(self.x, self.y, h, 0.0, 0.0, self.dx, self.dy, self.f)
        shrunk = [dict(states[0])] + [
            dict(s, x_min=s["xmin"] + self.dx / 100.0, y_min=s["ymin"] + self.dy / 100.0,
                 x_max=s["xmax"] - self.dx / 100.0, y_max=s["ymax"] - self.dy / 100.0)
            for s in states[1:]]
        self.states = shrunk
        oracle.set_chunk_state(self.x, self.y, shrunk, self.f)
        for name in ("density", "energy0", "energy"):
            oracle.halo_field(self.x, self.y, h, 2, self.f[name])
        self.f["energy"][:] = self.f["energy0"]
