"""ctypes binding of libteahip.so (the C ABI in include/tea_hip.h).

Loading fails with a clear error when the library has not been built; there
is no fallback implementation.
"""

import ctypes
import os
import re

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libteahip.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "tea_hip.h")
TEALEAF_BIN = os.path.join(PKG_DIR, "bin", "tealeaf")

# Field ids (enum in tea_hip.h, order of TeaLeaf/chunk.h:39-66).
FIELDS = [
    "density0", "density", "energy0", "energy", "u", "u0", "p", "r", "mi",
    "w", "kx", "ky", "sd", "cell_x", "cell_y", "cell_dx", "cell_dy",
    "vertex_dx", "vertex_dy", "vertex_x", "vertex_y", "volume", "x_area",
    "y_area", "comms_buffer",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
# Halo slots (Settings::fields_to_exchange, TeaLeaf/shared.h:39-44).
HALO_SLOTS = ["density", "energy0", "energy1", "u", "p", "sd"]
CONDUCTIVITY, RECIP_CONDUCTIVITY = 1, 2
GEOMETRY = {"rectangle": 0, "circular": 1, "point": 2}
MP_FIELDS = {"p": 0, "r": 1, "w": 2, "kx": 3, "ky": 4}

_lib = None

_d = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_i = ctypes.c_int

_SIGNATURES = {
    "tea_hip_chunk_create": [_i, _i, _i, _i, ctypes.POINTER(_vp)],
    "tea_hip_chunk_destroy": [_vp],
    "tea_hip_field": [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t)],
    "tea_hip_field_write": [_vp, _i, _dp, ctypes.c_size_t],
    "tea_hip_field_read": [_vp, _i, _dp, ctypes.c_size_t],
    "tea_hip_synchronize": [_vp],
    "tea_hip_set_chunk_data": [_vp, _d, _d, _d, _d],
    "tea_hip_set_chunk_state": [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip],
    "tea_hip_store_energy": [_vp],
    "tea_hip_local_halos": [_vp, _i, _ip, _ip],
    "tea_hip_field_summary": [_vp, _dp, _dp, _dp, _dp],
    "tea_hip_cg_init": [_vp, _i, _d, _d, _dp],
    "tea_hip_cg_calc_w": [_vp, _dp],
    "tea_hip_cg_calc_ur": [_vp, _d, _dp],
    "tea_hip_cg_calc_p": [_vp, _d],
    "tea_hip_copy_u": [_vp],
    "tea_hip_calculate_residual": [_vp],
    "tea_hip_calculate_2norm": [_vp, _vp, _dp],
    "tea_hip_calculate_2norm_noalloc": [_i, _vp, _i, _i, _i, _vp, _vp],
    "tea_hip_finalise": [_vp],
    "tea_hip_pack_or_unpack": [_vp, _i, _i, _i, _vp, _dp],
    "tea_hip_jacobi_init": [_vp, _i, _d, _d],
    "tea_hip_jacobi_iterate": [_vp, _dp],
    "tea_hip_cheby_init": [_vp, _d],
    "tea_hip_cheby_iterate": [_vp, _d, _d],
    "tea_hip_ppcg_init": [_vp, _d],
    "tea_hip_ppcg_inner_iteration": [_vp, _d, _d],
    "tea_hip_mpcg_init": [_vp, _i, _d, _d, _dp, _dp],
    "tea_hip_mpcg_calc_w": [_vp, _d, _i, _d, _i, _dp, _dp],
    "tea_hip_mpcg_calc_r": [_vp, _d, _dp, _dp],
    "tea_hip_mpcg_refresh": [_vp, _d, _i, _dp, _dp, _dp],
    "tea_hip_mp_read": [_vp, _i, _fp, ctypes.c_size_t],
    "tea_hip_fcg_attach": [_vp, _fp, ctypes.c_size_t, _fp, _i, _i],
    "tea_hip_fcg_attach_sweeps": [_vp, _i, _i, _i],
    "tea_hip_fcg_levels": [_vp, _ip],
    "tea_hip_fcg_init": [_vp, _i, _d, _d, _dp, _dp],
    "tea_hip_fcg_precondition": [_vp, _i, _dp],
    "tea_hip_fcg_calc_r_precondition": [_vp, _d, _i, _dp, _dp],
    "tea_hip_fcg_calc_w": [_vp, _d, _i, _d, _dp, _dp],
    "tea_hip_fcg_read_z": [_vp, _fp, ctypes.c_size_t],
    "tea_hip_set_fused": [_vp, _i],
    "tea_hip_flush": [_vp],
    "tea_hip_timing_enable": [_vp, _i],
    "tea_hip_kernel_stats": [_vp, _i, ctypes.POINTER(ctypes.c_long), _dp],
    "tea_hip_timing_reset": [_vp],
    "tea_hip_num_kernels": [],
    "tea_hip_copy_bandwidth": [_i, ctypes.c_size_t, _i, _dp],
    "tea_hip_kernel_name": [_i],
    "tea_hip_last_error": [],
}


class TeaHipError(RuntimeError):
    pass


def header_symbols(path=HEADER_PATH):
    """Names of the functions declared in include/tea_hip.h."""
    text = open(path).read()
    return sorted(set(re.findall(r"\b(tea_(?:hip|nn)_\w+)\s*\(", text)))


def load(path=LIB_PATH):
    """Loads libteahip.so once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TeaHipError(
            f"{path} is not built; run `make -C exploringsycl_amd` "
            "(or __graft_entry__.build()) first")
    lib = ctypes.CDLL(path)
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.tea_hip_last_error.restype = ctypes.c_char_p
    lib.tea_hip_kernel_name.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(status, what=""):
    if status != 0:
        msg = load().tea_hip_last_error().decode()
        raise TeaHipError(f"{what} failed ({status}): {msg}")


def kernel_names():
    lib = load()
    return [lib.tea_hip_kernel_name(k).decode()
            for k in range(lib.tea_hip_num_kernels())]
