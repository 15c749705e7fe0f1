"""Loader for libigr_hip.so, the C ABI declared in include/igr.h.

The library is built in-tree (igr_sensitivities_amd/lib/) by
igr_sensitivities_amd.build.build().  There is no fallback: if the shared
library is missing or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libigr_hip.so")

IGR_OK, IGR_EINVAL, IGR_EDOMAIN, IGR_EHIP, IGR_ENOMEM = 0, 1, 2, 3, 4

_d, _i, _ll = C.c_double, C.c_int, C.c_longlong
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_vp = C.c_void_p


class BatchItem(C.Structure):
    """struct igr_batch_item of include/igr.h."""
    _fields_ = [("ctx", _vp), ("dt", _d), ("t_end", _d), ("gamma", _d), ("alpha", _d), ("sgam", _d),
                ("n_tan", _i), ("s0", _pd), ("s_out", _pd), ("vs_out", _pd), ("n_steps", _i), ("status", _i)]


# name -> (restype, argtypes); every exported symbol of include/igr.h
PROTOTYPES = {
    "igr_create": (_i, [C.POINTER(_vp), _i, _i, _i, _i, _d, _d]),
    "igr_destroy": (_i, [_vp]),
    "igr_last_error": (C.c_char_p, [_vp]),
    "igr_sizes": (_i, [_vp, C.POINTER(_ll), _pi]),
    "igr_set_option": (_i, [_vp, _i, _i]),
    "igr_get_option": (_i, [_vp, _i, _pi]),
    "igr_set_state": (_i, [_vp, _pd, _pd]),
    "igr_get_state": (_i, [_vp, _pd, _pd]),
    "igr_hyperbolic_rhs": (_i, [_vp, _d, _pd, _pd, _pd]),
    "igr_apply_sip": (_i, [_vp, _d, _pd, _pd, _pd]),
    "igr_igr_rhs": (_i, [_vp, _d, _pd, _pd]),
    "igr_sip_diagonal": (_i, [_vp, _d, _pd, _pd]),
    "igr_solve_elliptic": (_i, [_vp, _d, _i, _i]),
    "igr_solver_trace": (_i, [_vp, _i, _pd, _pd]),
    "igr_euler_rhs": (_i, [_vp, _d, _d, _i, _i, _pd]),
    "igr_run_forward": (_i, [_vp, _d, _d, _d, _d, _i, _i, _pi]),
    "igr_run_steps": (_i, [_vp, _i, _d, _d, _d, _i, _i]),
    "igr_time_kernel": (_i, [_vp, _i, _d, _d, _i, _pd]),
    "igr_objective": (_i, [_vp, _i, _pd, _d, _pd]),
    "igr_run_sensitivity": (_i, [_vp, _i, _pd, _d, _d, _d, _d, _i, _i, _i, _d, _pd, _pd, _pi]),
    "igr_gradient": (_i, [_vp, _i, _pd, _i, _d, _d, _d, _i, _i, _i, _pd, _pd, _pd, _pd]),
    "igr_hyperbolic_vjp": (_i, [_vp, _d, _pd, _pd, _pd, _pd, _pd, _pd]),
    "igr_apply_sip_vjp": (_i, [_vp, _d, _pd, _pd, _pd, _pd, _pd, _pd]),
    "igr_solve_elliptic_vjp": (_i, [_vp, _d, _i, _i, _pd, _pd, _pd, _pd, _pd]),
    "igr_batch_run_sensitivity": (_i, [C.POINTER(BatchItem), _i, _i, _i, _i]),
    "igr_run_forward_store": (_i, [_vp, _d, _d, _d, _d, _i, _i, _pi]),
    "igr_snapshot_count": (_i, [_vp, _pi]),
    "igr_get_snapshot": (_i, [_vp, _i, _pd, _pd]),
    "igr_set_snapshots": (_i, [_vp, _i, _pd]),
    "igr_run_adjoint_conservative": (_i, [_vp, _i, _pd, _i, _d, _d, _d, _d, _i, _i, _pd]),
    "igr_release_cache": (_i, []),
    "igr_build_ic": (_i, [_vp, _i, _d, _d, _pd, _i, _pd]),
    "igr_solve_elliptic_noalloc": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """The loaded library (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with igr_sensitivities_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
