"""MI355X-native IGR-regularized Euler time step, tangent and discrete adjoint.

A drop-in for the hot path of f-t-s/IGR_sensitivities: the shared library
libigr_hip.so (C ABI, include/igr.h) holds every kernel; this package only
loads it and mirrors the reference's Julia operator API on numpy arrays.
"""
from .core import (DGBasis, CartesianMesh2D, PeriodicMesh1D, Context, DomainError, IGRError,  # noqa: F401
                   SOLVERS, compute_hyperbolic_rhs, apply_sip_, compute_igr_rhs_, compute_sip_diagonal,
                   solve_elliptic_, compute_igr_euler_rhs_, run_forward_, run_forward,
                   enzyme_adjoint_ic, OBJECTIVES, run_sensitivity, run_sensitivity_batch,
                   run_forward_store, run_adjoint_conservative, Snapshots, release_cache, build_ic, IC_TYPES)
from ._lib import lib, LIB_PATH  # noqa: F401
