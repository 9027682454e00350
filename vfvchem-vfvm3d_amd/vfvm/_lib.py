"""ctypes binding of the C ABI declared in include/vfvm.h.

The shared library lib/libvfvm.so next to this package is required: there is
no fallback implementation, and loading fails loudly when it is absent.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libvfvm.so")

# status codes (include/vfvm.h, vfvm_status)
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_HIP = 2
ERR_TETRA_DEGENERATE_INPUT = 10
ERR_TETRA_POINTS_ON_BOUNDARY = 11
ERR_TETRA_SINGULAR_GRADIENT = 12
ERR_OPERATOR_NON_UNKNOWN = 20
ERR_OPERATOR_USE_APPLY_TIME = 21
ERR_OPERATOR_UNSUPPORTED = 22
ERR_PATTERN_MISMATCH = 23
ERR_LINEAR_NON_POSITIVE_TOLERANCE = 30
ERR_LINEAR_NON_POSITIVE_RESTART = 31
ERR_LINEAR_NON_POSITIVE_MAX_OUTER = 32
ERR_LINEAR_DID_NOT_CONVERGE = 33
ERR_LINEAR_NON_FINITE = 34
ERR_SOLVER_NON_POSITIVE_MAX_ITER = 40
ERR_SOLVER_NON_POSITIVE_TOLERANCE = 41
ERR_SOLVER_INVALID_DAMPING = 42
ERR_SOLVER_NON_POSITIVE_TIME_STEP = 43
ERR_SOLVER_NON_POSITIVE_NUM_STEPS = 44
ERR_STEADY_DID_NOT_CONVERGE = 45
ERR_STEADY_NON_FINITE_RESIDUAL = 46
ERR_TRANSIENT_DID_NOT_CONVERGE = 47
ERR_TRANSIENT_NON_FINITE_RESIDUAL = 48
ERR_WRITE_CSV = 50
ERR_WRITE_VTU = 51
ERR_VARIABLES_NON_POSITIVE_SCALE = 60

STATUS_NAMES = {v: k for k, v in globals().items() if k.startswith("ERR_") or k == "OK"}


class VFVChemError(RuntimeError):
    """Error raised for a non-zero status; mirrors VFVChemError (src/base/error.rs:5)."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = STATUS_NAMES.get(code, str(code))
        super().__init__(f"{self.kind}: {message}")


class MeshInfo(C.Structure):
    _fields_ = [("num_pts", C.c_int64), ("num_nds", C.c_int64), ("num_tets", C.c_int64),
                ("num_edges", C.c_int64), ("num_subfaces", C.c_int64), ("num_bnd", C.c_int32),
                ("max_degree", C.c_int32), ("num_grad_coeff", C.c_int64), ("num_dom", C.c_int32), ("num_itf", C.c_int32),
                ("num_itf_nodes", C.c_int64)]


class OpStruct(C.Structure):
    _fields_ = [("kind", C.c_int32), ("geom_id", C.c_int32), ("unk_id", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("c", C.c_int32), ("value", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("restart", C.c_int32), ("max_outer", C.c_int32), ("precision", C.c_int32),
                ("precond", C.c_int32), ("inner_tol", C.c_double), ("max_refine", C.c_int32), ("ortho", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("restarts", C.c_int32), ("refinements", C.c_int32), ("status", C.c_int32),
                ("rel_residual", C.c_double), ("seconds", C.c_double), ("setup_seconds", C.c_double)]


class AdaptiveOpts(C.Structure):
    _fields_ = [("t_end", C.c_double), ("dt_init", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("rtol", C.c_double), ("atol", C.c_double), ("safety", C.c_double), ("fac_min", C.c_double),
                ("fac_max", C.c_double), ("fac_fail", C.c_double), ("max_reject", C.c_int32), ("max_iter", C.c_int32),
                ("tol", C.c_double), ("damp", C.c_double), ("fixed", C.c_int32), ("num_out", C.c_int32),
                ("out_times", C.POINTER(C.c_double)), ("num_exclude", C.c_int32), ("reserved", C.c_int32),
                ("scl_exclude", C.POINTER(C.c_int32))]


class AdaptiveStats(C.Structure):
    _fields_ = [("accepted", C.c_int32), ("rejected_error", C.c_int32), ("rejected_picard", C.c_int32),
                ("picard_iterations", C.c_int32), ("linear_iterations", C.c_int32), ("outputs", C.c_int32),
                ("t_final", C.c_double), ("time_assembly", C.c_double), ("time_solve", C.c_double)]


STEP_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_double)
UPDATE_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_double)


class DriverStats(C.Structure):
    _fields_ = [("picard_iterations", C.c_int32), ("linear_iterations", C.c_int32), ("residual", C.c_double),
                ("time_assembly", C.c_double), ("time_solve", C.c_double)]


class AdjointStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("checkpoints", C.c_int32), ("forward_iterations", C.c_int32),
                ("recompute_iterations", C.c_int32), ("adjoint_iterations", C.c_int32), ("recomputed_steps", C.c_int32),
                ("replay_mismatches", C.c_int32), ("reserved", C.c_int32),
                ("forward_seconds", C.c_double), ("backward_seconds", C.c_double), ("recompute_seconds", C.c_double),
                ("adjoint_solve_seconds", C.c_double), ("setup_seconds", C.c_double)]


class AdjointWrt(C.Structure):
    _fields_ = [("kind", C.c_int32), ("id", C.c_int32), ("grad", C.POINTER(C.c_double))]


ADJOINT_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double)


class FlowAdjointOpts(C.Structure):
    _fields_ = [("inner", SolveOpts), ("tol", C.c_double), ("max_iter", C.c_int32), ("restart", C.c_int32),
                ("on_state", ADJOINT_HOOK), ("user", C.c_void_p)]


class FlowAdjointStats(C.Structure):
    _fields_ = [("base", AdjointStats), ("picard_iterations", C.c_int32), ("recompute_picard_iterations", C.c_int32),
                ("outer_iterations", C.c_int32), ("reserved", C.c_int32), ("objective", C.c_double),
                ("max_outer_residual", C.c_double)]


_P = C.c_void_p
_I = C.c_int32
_L = C.c_int64
_D = C.c_double
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int32)
_PL = C.POINTER(C.c_int64)

# (name, restype, argtypes) for every entry point of include/vfvm.h
SIGNATURES = [
    ("vfvm_last_error", C.c_char_p, []),
    ("vfvm_version", C.c_char_p, []),
    ("vfvm_mesh_lattice", _I, [_I, C.c_uint64, _D, C.POINTER(_P)]),
    ("vfvm_mesh_from_tets", _I, [_L, _L, _PD, _L, _PI, _I, _PL, _PI, C.POINTER(_P)]),
    ("vfvm_mesh_lattice_dev", _I, [_I, C.c_uint64, _D, _I, C.POINTER(_P)]),
    ("vfvm_mesh_lattice_box", _I, [_I, C.c_uint64, _D, _PI, _PI, _I, C.POINTER(_P)]),
    ("vfvm_mesh_join", _I, [_I, C.POINTER(_P), _I, _PI, C.POINTER(_P)]),
    ("vfvm_mesh_from_tets_dev", _I, [_L, _L, _PD, _L, _PI, _I, _PL, _PI, _I, C.POINTER(_P)]),
    ("vfvm_mesh_info_get", _I, [_P, C.POINTER(MeshInfo)]),
    ("vfvm_mesh_build_grad_coeff", _I, [_P]),
    ("vfvm_mesh_build_grad_coeff_dev", _I, [_P, _I]),
    ("vfvm_mesh_get", _I, [_P, _I, _P]),
    ("vfvm_write_csv", _I, [_P, _I, _I, _PD, C.c_char_p, C.c_char_p, _I, _I]),
    ("vfvm_write_vtu", _I, [_P, _I, _I, _PD, C.c_char_p, C.c_char_p, _I, _I]),
    ("vfvm_mesh_destroy", None, [_P]),
    ("vfvm_create", _I, [_I, C.POINTER(_P)]),
    ("vfvm_destroy", None, [_P]),
    ("vfvm_mesh_upload", _I, [_P, _P]),
    ("vfvm_scldom_add", _I, [_P, _I, _D, _PD, _PI]),
    ("vfvm_scldom_set", _I, [_P, _I, _PD]),
    ("vfvm_scldom_get", _I, [_P, _I, _PD]),
    ("vfvm_scldom_update_prev", _I, [_P, _I]),
    ("vfvm_vecdom_add", _I, [_P, _I, _PD, _PD, _PI]),
    ("vfvm_vecdom_set", _I, [_P, _I, _PD]),
    ("vfvm_vecdom_get", _I, [_P, _I, _PD]),
    ("vfvm_sclbnd_add", _I, [_P, _I, _D, _PD, _PI]),
    ("vfvm_sclbnd_set", _I, [_P, _I, _PD]),
    ("vfvm_vecbnd_set", _I, [_P, _I, _PD]),
    ("vfvm_sclflw_add", _I, [_P, _I, _I, _I, _I, _I, _PI]),
    ("vfvm_vecbnd_add", _I, [_P, _I, _PD, _PD, _PI]),
    ("vfvm_update_flow", _I, [_P, _I]),
    ("vfvm_sclflw_get", _I, [_P, _I, _I, _PD]),
    ("vfvm_pattern_build", _I, [_P, C.POINTER(OpStruct), _I, _PL]),
    ("vfvm_pattern_get", _I, [_P, _PI, _PI]),
    ("vfvm_system_info", _I, [_P, _PL, _PL]),
    ("vfvm_set_update_hook", _I, [_P, UPDATE_HOOK, _P]),
    ("vfvm_assemble", _I, [_P, C.POINTER(OpStruct), _I, _D]),
    ("vfvm_matrix_get", _I, [_P, _PD, _PD]),
    ("vfvm_solve_opts_default", None, [C.POINTER(SolveOpts)]),
    ("vfvm_solve", _I, [_P, C.POINTER(SolveOpts), _PD, C.POINTER(SolveStats)]),
    ("vfvm_solve_device", _I, [_P, C.POINTER(SolveOpts), _P, C.POINTER(SolveStats)]),
    ("vfvm_residual", _I, [_P, _PD, _PD]),
    ("vfvm_precond_apply", _I, [_P, C.POINTER(SolveOpts), _PD, _PD]),
    ("vfvm_ilu_order", _I, [_P, _PI, _PI, _PI]),
    ("vfvm_ilu_tiles", _I, [_P, _PI, _PI, _PI]),
    ("vfvm_ilu_wait_limit", _I, [_P, C.c_int32, C.POINTER(C.c_uint64)]),
    ("vfvm_steady_solve", _I, [_P, C.POINTER(OpStruct), _I, C.POINTER(SolveOpts), _I, _D, _D, C.POINTER(DriverStats)]),
    ("vfvm_adaptive_opts_default", None, [_D, _D, C.POINTER(AdaptiveOpts)]),
    ("vfvm_transient_adaptive", _I, [_P, C.POINTER(OpStruct), _I, C.POINTER(SolveOpts), C.POINTER(AdaptiveOpts), STEP_HOOK,
                                     _P, _PD, _I, C.POINTER(AdaptiveStats)]),
    ("vfvm_transient_solve", _I, [_P, C.POINTER(OpStruct), _I, C.POINTER(SolveOpts), _D, _I, _I, _D, _D, C.POINTER(DriverStats)]),
    ("vfvm_adjoint_forward", _I, [_P, C.POINTER(OpStruct), _I, C.POINTER(SolveOpts), _D, _I, _I, C.POINTER(AdjointStats)]),
    ("vfvm_adjoint_backward", _I, [_P, C.POINTER(SolveOpts), _PD, _I, C.POINTER(AdjointWrt), C.POINTER(AdjointStats)]),
    ("vfvm_vecdom_func_affine", _I, [_P, _I, _I, _PD, _PD]),
    ("vfvm_flow_adjoint_forward", _I, [_P, C.POINTER(OpStruct), _I, C.POINTER(SolveOpts), _D, _I, _I, _D, _D, _I, _PD, _PD,
                                       C.POINTER(FlowAdjointStats)]),
    ("vfvm_flow_adjoint_backward", _I, [_P, C.POINTER(FlowAdjointOpts), _PD, _PD, _I, C.POINTER(AdjointWrt),
                                        C.POINTER(FlowAdjointStats)]),
    ("vfvm_flow_residual_eval", _I, [_P, C.POINTER(OpStruct), _I, _D, _PD, _PD, _PD]),
    ("vfvm_flow_jt_apply", _I, [_P, C.POINTER(OpStruct), _I, _D, _PD, _PD, _PD, _PD]),
    ("vfvm_set_scldom_scale", _I, [_P, _I, _D]),
    ("vfvm_set_vecdom_scale", _I, [_P, _I, _D]),
    ("vfvm_set_scale_off", _I, [_P, _I]),
    ("vfvm_linear_update_sparse", _I, [_P, _L, _L, _PL, _PL, _PL]),
    ("vfvm_linear_build_matrix", _I, [_P, _L, _PD]),
    ("vfvm_linear_build_matrix_device", _I, [_P, _L, _P]),
    ("vfvm_linear_solve_matrix", _I, [_P, C.POINTER(SolveOpts), _PD, _PD, _PD, C.POINTER(SolveStats)]),
    ("vfvm_linear_solve_matrix_device", _I, [_P, C.POINTER(SolveOpts), _P, _P, C.POINTER(SolveStats)]),
    ("vfvm_solve_device_noalloc", _I, [_P, C.POINTER(SolveOpts), _P, C.POINTER(SolveStats)]),
    ("vfvm_timing_enable", _I, [_P, _I]),
    ("vfvm_timing_get", _I, [_P, _I, _PL, _PD, _PD]),
    ("vfvm_timing_reset", _I, [_P]),
]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"vfvm: native library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(code: int) -> None:
    if code != OK:
        msg = lib.vfvm_last_error()
        raise VFVChemError(code, msg.decode() if msg else "")


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))
