"""ctypes view of the C ABI declared in include/lbm_core.h.

The HIP core is a shared library built in-tree (lb_amd/liblbm_core_hip.so).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
CORE_LIB = os.path.join(HERE, "liblbm_core_hip.so")
HOST_LIB = os.path.join(HERE, "liblbm_host.so")

Q = 27

# status codes
LBM_OK, LBM_E_ARG, LBM_E_HIP, LBM_E_OOM, LBM_E_NAN, LBM_E_STATE, LBM_E_UNSUPPORTED = range(7)
STATUS_NAMES = {0: "LBM_OK", 1: "LBM_E_ARG", 2: "LBM_E_HIP", 3: "LBM_E_OOM", 4: "LBM_E_NAN",
                5: "LBM_E_STATE", 6: "LBM_E_UNSUPPORTED"}

COLLISION_SRT, COLLISION_CUM2015, COLLISION_CUM2017, COLLISION_COMBINED = 0, 1, 2, 3
CLOSURE_NONE, CLOSURE_SMAGORINSKY, CLOSURE_MLP = 0, 1, 2
STORAGE_F64, STORAGE_F32, STORAGE_F16DEV, STORAGE_BF16DEV = 0, 1, 2, 3
STORAGE_F16DEV_SR, STORAGE_BF16DEV_SR = 4, 5   # stochastic rounding in the step's stores
ARITH_LITERAL, ARITH_FAST = 0, 1
COMPAT_CREV20, COMPAT_F32CONST, COMPAT_LIM210, COMPAT_TRACE17 = 0x1, 0x2, 0x4, 0x8
COMPAT_INLETRHO, COMPAT_EQ44, COMPAT_OMEGA4 = 0x10, 0x20, 0x40
COMPAT_REFERENCE = (COMPAT_CREV20 | COMPAT_F32CONST | COMPAT_LIM210 | COMPAT_TRACE17 | COMPAT_INLETRHO |
                    COMPAT_EQ44 | COMPAT_OMEGA4)
BC_WALL, BC_INLET_MOVING_WALL, BC_INLET_EQUILIBRIUM, BC_OUTLET_INTERPOLATED, BC_PERIODIC_GHOST = range(5)
BC_OUTLET_EQUILIBRIUM, BC_OUTLET_NEIGHBOUR_EQ, BC_OUTLET_NEIGHBOUR_EQ_RF, BC_INLET_ZOU_HE = 5, 6, 7, 8

STORAGE_BYTES = {STORAGE_F64: 8, STORAGE_F32: 4, STORAGE_F16DEV: 2, STORAGE_BF16DEV: 2, STORAGE_F16DEV_SR: 2,
                 STORAGE_BF16DEV_SR: 2}
STORAGE_NAMES = {STORAGE_F64: "f64", STORAGE_F32: "f32", STORAGE_F16DEV: "f16dev", STORAGE_BF16DEV: "bf16dev",
                 STORAGE_F16DEV_SR: "f16dev_sr", STORAGE_BF16DEV_SR: "bf16dev_sr"}


class BcNode(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("nx", ctypes.c_float), ("ny", ctypes.c_float), ("nz", ctypes.c_float),
                ("ux", ctypes.c_float), ("uy", ctypes.c_float), ("uz", ctypes.c_float),
                ("rho", ctypes.c_float), ("per_idx", ctypes.c_int32), ("regular", ctypes.c_int32)]


class Desc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("periodic", ctypes.c_int32 * 3),
                ("collision", ctypes.c_int32), ("closure", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("arith", ctypes.c_int32), ("compat", ctypes.c_uint32), ("n_bc", ctypes.c_int32),
                ("tau", ctypes.c_double), ("nu", ctypes.c_double), ("c_les", ctypes.c_double),
                ("rho0", ctypes.c_double),
                ("flags", ctypes.POINTER(ctypes.c_int32)), ("bc", ctypes.POINTER(BcNode)),
                ("outlet_sponge", ctypes.c_int32), ("sponge_low", ctypes.c_float), ("sponge_high", ctypes.c_float),
                ("wall_mesh", ctypes.c_int32)]


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every entry point declared in include/lbm_core.h
CORE_SYMBOLS = [
    ("lbm_create", ctypes.c_int, [ctypes.POINTER(Desc), ctypes.POINTER(_vp)]),
    ("lbm_init_equilibrium", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    ("lbm_set_populations", ctypes.c_int, [_vp, _dp]),
    ("lbm_step", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("lbm_sync", ctypes.c_int, [_vp]),
    ("lbm_get_macroscopic", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    ("lbm_get_populations", ctypes.c_int, [_vp, _dp]),
    ("lbm_probe", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp]),
    ("lbm_probe_record", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("lbm_probe_fetch", ctypes.c_int, [_vp, _dp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    ("lbm_get_omega", ctypes.c_int, [_vp, _dp]),
    ("lbm_conservation", ctypes.c_int, [_vp, _dp]),
    ("lbm_step_diagnostics", ctypes.c_int, [_vp, _dp]),
    ("lbm_error_quadratic", ctypes.c_int, [_vp, _dp]),
    ("lbm_get_population", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    ("lbm_snapshot_begin", ctypes.c_int, [_vp]),
    ("lbm_snapshot_wait", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int64)]),
    ("lbm_step_diagnostics_noalloc", ctypes.c_int, [_vp, _dp]),
    ("lbm_moment_add", ctypes.c_int, [_vp]),
    ("lbm_step_moment_add", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("lbm_moment_avg", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int64)]),
    ("lbm_moment_avg_reset", ctypes.c_int, [_vp]),
    ("lbm_step_count", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    ("lbm_last_kernel_ms", ctypes.c_int, [_vp, _dp]),
    ("lbm_device_bytes", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    ("lbm_bandwidth_probe", ctypes.c_int, [ctypes.c_size_t, ctypes.c_int32, _dp, _dp]),
    ("lbm_device_info", ctypes.c_int, [ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p, ctypes.c_size_t]),
    ("lbm_voxelize", ctypes.c_int, [_dp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, _vp,
                                    ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                    ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.POINTER(ctypes.c_int32)]),
    ("lbm_status_str", ctypes.c_char_p, [ctypes.c_int]),
    ("lbm_destroy", None, [_vp]),
]

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)

MLP_FEATURES = 12
MLP_PARAMS = 5057

# (name, restype, argtypes) of every entry point declared in include/lbm_closure.h
CLOSURE_SYMBOLS = [
    ("lbm_mlp_create", ctypes.c_int, [_fp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("lbm_mlp_set_data", ctypes.c_int, [_vp, _fp, _fp, _fp, ctypes.c_int64]),
    ("lbm_mlp_gradient", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, _dp,
                                        _fp]),
    ("lbm_mlp_train", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_uint64, _dp]),
    ("lbm_mlp_forward", ctypes.c_int, [_vp, _fp, ctypes.c_int64, ctypes.c_float, _fp]),
    ("lbm_mlp_get_params", ctypes.c_int, [_vp, _fp]),
    ("lbm_mlp_set_params", ctypes.c_int, [_vp, _fp]),
    ("lbm_mlp_destroy", None, [_vp]),
    ("lbm_set_closure", ctypes.c_int, [_vp, _fp, _fp]),
    ("lbm_set_closure_refresh", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("lbm_get_closure_features", ctypes.c_int, [_vp, _fp, _ip, _dp]),
    ("lbm_filter", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
]

_core = None


def load_core() -> ctypes.CDLL:
    """Load the HIP core.  Raises if it has not been built."""
    global _core
    if _core is None:
        if not os.path.exists(CORE_LIB):
            raise RuntimeError(f"HIP core {CORE_LIB} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(CORE_LIB)
        for name, res, args in CORE_SYMBOLS + CLOSURE_SYMBOLS:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _core = lib
    return _core


def device_info():
    """(count, description) of the visible HIP devices; (0, reason) when none."""
    lib = load_core()
    n = ctypes.c_int32(0)
    buf = ctypes.create_string_buffer(256)
    st = lib.lbm_device_info(ctypes.byref(n), buf, 256)
    if st != LBM_OK:
        return 0, STATUS_NAMES.get(st, str(st))
    return n.value, buf.value.decode(errors="replace")


def ptr(a, ctype=ctypes.c_double):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))
