"""ctypes view of the C ABI in include/cbm_hip.h and the loader of libcbm_hip.so.

The structures here are byte-for-byte the C structs of the header; the
library is loaded from the package's ``lib/`` folder (built in-tree by
``csrc/build.sh``).  There is no fallback: if the shared library is missing
the import of :func:`lib` raises.
"""
import ctypes as C
import os

# ---------------------------------------------------------------- constants
OK = 0
ERR_CONFIG, ERR_CAPACITY, ERR_LIST_OVERFLOW, ERR_COMPILE, ERR_HIP, ERR_ARG, ERR_NONFINITE = -1, -2, -3, -4, -5, -6, -7
ERRORS = {
    ERR_CONFIG: "CBM_ERR_CONFIG", ERR_CAPACITY: "CBM_ERR_CAPACITY", ERR_LIST_OVERFLOW: "CBM_ERR_LIST_OVERFLOW",
    ERR_COMPILE: "CBM_ERR_COMPILE", ERR_HIP: "CBM_ERR_HIP", ERR_ARG: "CBM_ERR_ARG", ERR_NONFINITE: "CBM_ERR_NONFINITE",
}

EULER, HEUN, RK4, EM, EULER_HEUN = 0, 1, 2, 3, 4
FULL, CELL_LINKED, VERLET_DISPLACEMENT, VERLET_TIME, CLVD = 0, 1, 2, 3, 4
LAW_NONE, LAW_LINEAR, LAW_REP_ADH, LAW_AGGREGATION, LAW_CUSTOM, LAW_SOFT_SPHERES, LAW_RODS = 0, 1, 2, 3, 4, 5, 6
RULE_NONE, RULE_GROWTH_DIVISION, RULE_PARITY, RULE_DEATH, RULE_CUSTOM, RULE_RODS_GROWTH = 0, 1, 2, 3, 4, 5
MEDIUM_FTCS, MEDIUM_DGADI = 0, 1
BORDER_MIRROR, BORDER_DIRICHLET, BORDER_COPY = 0, 1, 2

F_X, F_Y, F_Z, F_ID, F_FX, F_FY, F_FZ, F_VOL, F_GROWTH, F_VBIRTH, F_LF = range(11)
F_VX, F_VY, F_VZ, F_RADIUS, F_MASS, F_FATE = range(11, 17)
F_U0, F_U1, F_U2, F_U3 = range(17, 21)   # user agent parameters (RULE_CUSTOM)
F_THETA, F_OMEGA, F_DIAMETER, F_LENGTH, F_TORQUE, F_PRESSURE = range(21, 27)   # rods
F_U4, F_U5, F_U6, F_U7 = range(27, 31)   # user agent parameters U4..U7
F_LTARGET = 31                            # rods with RULE_RODS_GROWTH
USER_PARAMS = 8


def user_field(k):
    """Field id of user agent parameter U_k (0 <= k < 8)."""
    if not 0 <= k < USER_PARAMS:
        raise ValueError("user agent parameters are U0..U7")
    return F_U0 + k if k < 4 else F_U4 + (k - 4)
INT_FIELDS = (F_ID, F_LF, F_FATE)


class ModelDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int32), ("integrator", C.c_int32), ("neighbors", C.c_int32), ("law", C.c_int32),
        ("rule", C.c_int32), ("periodic", C.c_int32), ("lagged", C.c_int32), ("max_neighbors", C.c_int32),
        ("dt", C.c_double), ("cell_edge", C.c_double), ("skin", C.c_double), ("rebuild_disp", C.c_double),
        ("dt_recompute", C.c_double), ("drift", C.c_double), ("noise", C.c_double),
        ("rule_p", C.c_double * 4), ("seed", C.c_uint64),
        ("medium_n", C.c_int32 * 3), ("medium_alg", C.c_int32),
        ("medium_D", C.c_double), ("medium_consume", C.c_double), ("medium_c0", C.c_double),
        ("user_de", C.c_int32), ("walk", C.c_int32),
        ("medium_border", C.c_int32 * 6), ("medium_border_value", C.c_double * 6),
    ]


class SimDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("n_max", C.c_int64),
        ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3),
        ("p", C.c_double * 8),
    ]


class LossDesc(C.Structure):
    _fields_ = [
        ("n_steps", C.c_int64), ("checkpoint_every", C.c_int64),
        ("n_saves", C.c_int32), ("pad", C.c_int32),
        ("save_steps", C.POINTER(C.c_int64)), ("target", C.POINTER(C.c_double)),
    ]


# Every symbol include/cbm_hip.h declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "cbm_create": (C.c_int, [C.POINTER(_P), C.POINTER(ModelDesc), C.c_int32, C.POINTER(SimDesc)]),
    "cbm_destroy": (None, [_P]),
    "cbm_last_error": (C.c_char_p, [_P]),
    "cbm_version": (C.c_char_p, []),
    "cbm_compile": (C.c_int, [_P, C.c_char_p]),
    "cbm_compile_model": (C.c_int, [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]),
    "cbm_upload": (C.c_int, [_P, C.c_int32, _P, C.c_int64, C.c_int32]),
    "cbm_download": (C.c_int, [_P, C.c_int32, _P, C.c_int64, C.c_int32]),
    "cbm_set_n": (C.c_int, [_P, C.c_int32, C.c_int64]),
    "cbm_upload_medium": (C.c_int, [_P, _P, C.c_int64]),
    "cbm_download_medium": (C.c_int, [_P, _P, C.c_int64]),
    "cbm_snapshot_config": (C.c_int, [_P, _P, C.c_int32, C.c_int32]),
    "cbm_snapshot_take": (C.c_int, [_P]),
    "cbm_step_save": (C.c_int, [_P, C.c_int64, C.c_int64]),
    "cbm_snapshot_count": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "cbm_snapshot_get": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, C.c_int64, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double)]),
    "cbm_snapshot_get_medium": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "cbm_neighbors": (C.c_int, [_P]),
    "cbm_step": (C.c_int, [_P, C.c_int64]),
    "cbm_step_graph": (C.c_int, [_P, C.c_int64, C.c_int64]),
    "cbm_sync": (C.c_int, [_P]),
    "cbm_get_n": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64)]),
    "cbm_get_t": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "cbm_get_id_max": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64)]),
    "cbm_get_rebuilds": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "cbm_get_bricks": (C.c_int, [_P, _P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "cbm_get_sub_bricks": (C.c_int, [_P, _P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "cbm_get_neighbor_lists": (C.c_int, [_P, C.c_int32, C.c_double, _P, _P, C.c_int64]),
    "cbm_count_neighbors": (C.c_int, [_P, C.c_int32, C.c_double, _P]),
    "cbm_count_neighbors_noalloc": (C.c_int, [_P, C.c_int32, C.c_double, _P]),
    "cbm_compile_check": (C.c_int, [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, _P, C.c_int64]),
    "cbm_adjoint": (C.c_int, [_P, C.POINTER(LossDesc), _P, _P]),
    "cbm_set_params": (C.c_int, [_P, C.c_int32, _P, C.c_int32]),
    "cbm_get_params": (C.c_int, [_P, C.c_int32, _P, C.c_int32]),
    "cbm_compile_model_scope": (C.c_int, [_P, C.c_char_p]),
    "cbm_compile_model_scope_check": (C.c_int, [C.c_char_p, _P, C.c_int64]),
    "cbm_autotune": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int32)]),
    "cbm_set_block": (C.c_int, [_P, C.c_int32]),
    "cbm_profile": (C.c_int, [_P, C.c_int64, _P, _P, C.c_int32]),
    "cbm_last_timing": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
}

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libcbm_hip.so")

_lib = None


def lib():
    """Load libcbm_hip.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with cellbasedmodels.jl_amd/csrc/build.sh")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


class CBMError(RuntimeError):
    """A negative status from the library (the reference raises `error()`)."""

    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(code, ctx=None):
    if code != OK:
        msg = lib().cbm_last_error(ctx).decode() if ctx else ""
        raise CBMError(code, msg)
