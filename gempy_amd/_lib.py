"""ctypes binding of libgkhip.so (declarations in include/gk.h).

The library is loaded from the package directory only; there is no fallback:
if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgkhip.so"

GK_OK = 0
GK_ERR = {-1: "GK_ERR_ARG", -2: "GK_ERR_HIP", -3: "GK_ERR_NOT_SPD", -4: "GK_ERR_SINGULAR",
          -5: "GK_ERR_NOT_CONVERGED", -6: "GK_ERR_UNSUPPORTED"}
GK_F32, GK_F64 = 0, 1
GK_POINTS_EXPLICIT, GK_POINTS_LATTICE = 0, 1
GK_SOLVE_FP64_LU, GK_SOLVE_FP32_REFINE, GK_SOLVE_FP64_CHOL = 0, 1, 2
GK_REL_ERODE, GK_REL_ONLAP = 1, 2
GK_MAX_THRESHOLDS = 64
GK_BLOCK = 256
GK_CHOL_PLANES = 5
GK_SPAI_WITH_FACTOR, GK_SPAI_NO_FALLBACK, GK_SPAI_SANDWICH = 1, 2, 4
GK_SPAI_K, GK_SPAI_FEATURES, GK_SPAI_HIDDEN = 16, 8, 32
GK_LOCAL_NEIGHBOURS = 32
GK_EVAL_SCRATCH_BYTES = 32 << 20

P = C.c_void_p


class GkSystem(C.Structure):
    _fields_ = [
        ("n_ori", C.c_int64), ("n_rest", C.c_int64), ("n_surf", C.c_int64),
        ("uni_degree", C.c_int32), ("n_fault", C.c_int32),
        ("range", C.c_double), ("c_o", C.c_double), ("i_res", C.c_double), ("gi_res", C.c_double),
        ("ori_xyz", P), ("ori_grad", P), ("ori_nugget", P),
        ("rest_xyz", P), ("rest_surf", P), ("rest_nugget", P), ("ref_xyz", P),
        ("fault_rest", P), ("center", C.c_double * 3),
    ]


class GkEvalDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("precision", C.c_int32), ("n_points", C.c_int64),
        ("xyz", P), ("origin", C.c_double * 3), ("spacing", C.c_double * 3), ("res", C.c_int64 * 3),
        ("fault_values", P), ("scalar", P), ("gradient", P), ("ids", P),
        ("n_thresholds", C.c_int32), ("thresholds", C.c_double * GK_MAX_THRESHOLDS),
    ]


class GkSolveInfo(C.Structure):
    _fields_ = [
        ("form", C.c_int32), ("iterations", C.c_int32), ("fallback", C.c_int32), ("info", C.c_int32),
        ("relres", C.c_double), ("t_assemble_ms", C.c_double), ("t_factor_ms", C.c_double),
        ("t_refine_ms", C.c_double), ("shift", C.c_double),
    ]


# name -> (restype, argtypes); every symbol declared in include/gk.h
SIGNATURES = {
    "gk_version": (C.c_int, []),
    "gk_last_error": (C.c_char_p, []),
    "gk_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "gk_destroy": (None, [P]),
    "gk_system_n": (C.c_int64, [C.POINTER(GkSystem)]),
    "gk_system_padded": (C.c_int64, [C.POINTER(GkSystem)]),
    "gk_assemble": (C.c_int, [P, C.POINTER(GkSystem), C.c_int32, P, C.c_int64, P, P, P]),
    "gk_assemble_drift": (C.c_int, [P, C.POINTER(GkSystem), P, P]),
    "gk_matvec": (C.c_int, [P, C.POINTER(GkSystem), C.c_int32, P, P, P]),
    "gk_cholesky_f32": (C.c_int, [P, P, C.c_int64, C.c_int64, P, C.POINTER(C.c_int32), P]),
    "gk_cholesky_solve_f32": (C.c_int, [P, P, P, C.c_int64, P, P]),
    "gk_cholesky_f64": (C.c_int, [P, P, C.c_int64, C.c_int64, P, C.POINTER(C.c_int32), P]),
    "gk_cholesky_solve_f64": (C.c_int, [P, P, P, C.c_int64, P, P]),
    "gk_lu_f64": (C.c_int, [P, P, C.c_int64, C.c_int64, P, C.POINTER(C.c_int32), P]),
    "gk_lu_solve_f64": (C.c_int, [P, P, P, C.c_int64, P, P]),
    "gk_gemm_f32": (C.c_int, [P, C.c_int64, C.c_int64, C.c_int64, P, C.c_int64, P, C.c_int64,
                              C.c_int32, C.c_float, P, C.c_int64, C.c_int32, P]),
    "gk_gemm_f64": (C.c_int, [P, C.c_int64, C.c_int64, C.c_int64, P, C.c_int64, P, C.c_int64,
                              C.c_int32, C.c_double, P, C.c_int64, C.c_int32, P]),
    "gk_solve": (C.c_int, [P, C.POINTER(GkSystem), C.c_int32, C.c_double, C.c_int32, P,
                           C.POINTER(GkSolveInfo), P]),
    "gk_solve_x0": (C.c_int, [P, C.POINTER(GkSystem), C.c_int32, C.c_double, C.c_int32, P, P,
                              C.POINTER(GkSolveInfo), P]),
    "gk_eval": (C.c_int, [P, C.POINTER(GkSystem), P, C.POINTER(GkEvalDesc), P]),
    "gk_eval_noalloc": (C.c_int, [P, C.POINTER(GkSystem), P, C.POINTER(GkEvalDesc), P]),
    "gk_solve_rhs": (C.c_int, [P, C.POINTER(GkSystem), P, P, C.POINTER(GkSolveInfo), P]),
    "gk_adjoint": (C.c_int, [P, C.POINTER(GkSystem), P, C.c_int64, P, P, P, P, P, P, P]),
    "gk_adjoint_full": (C.c_int, [P, C.POINTER(GkSystem), P, C.c_int64, P, P, P, P, P, P, P, P]),
    "gk_eval_sources": (C.c_int, [P, C.c_double, C.c_double, C.POINTER(C.c_double), C.c_int64, P, P,
                                  C.c_int64, P, P, P, P]),
    "gk_spai_nparams": (C.c_int, []),
    "gk_knn": (C.c_int, [P, C.POINTER(GkSystem), P, P, P]),
    "gk_spai_features": (C.c_int, [P, C.POINTER(GkSystem), C.c_int64, P, P, P, P, P]),
    "gk_spai_local_blocks": (C.c_int, [P, C.POINTER(GkSystem), C.c_int64, P, P, P, P, P]),
    "gk_mlp_forward": (C.c_int, [P, C.c_int64, P, P, P, P, P]),
    "gk_mlp_backward": (C.c_int, [P, C.c_int64, P, P, P, P, P, P]),
    "gk_spai_loss": (C.c_int, [P, C.POINTER(GkSystem), C.c_int64, P, P, P, P, P, P, P]),
    "gk_adam": (C.c_int, [P, P, P, P, P, C.c_int32, C.c_float, C.c_int32, C.c_float, P]),
    "gk_spai_apply": (C.c_int, [P, C.POINTER(GkSystem), P, P, P, P, P]),
    "gk_diag_scale": (C.c_int, [P, C.POINTER(GkSystem), P, P]),
    "gk_spai_local_inverse": (C.c_int, [P, C.POINTER(GkSystem), P, P, P, P]),
    "gk_solve_spai": (C.c_int, [P, C.POINTER(GkSystem), P, P, C.c_double, C.c_int32, C.c_int32, C.c_int32, P,
                                C.POINTER(GkSolveInfo), P]),
    "gk_profile": (C.c_int, [P, C.c_int32]),
    "gk_profile_read": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
    "gk_combine_lith": (C.c_int, [P, C.c_int32, P, C.c_int64, C.POINTER(C.c_int32), P, P]),
    "gk_marching_cubes": (C.c_int, [P, P, C.POINTER(C.c_int64), P, C.c_double, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), P, C.c_int64, P, C.c_int64, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), P]),
    "gk_condition_number": (C.c_int, [P, C.POINTER(GkSystem), C.c_int32, C.c_double, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_int32), P]),
    "gk_test_knn_local": (C.c_int, [P, C.POINTER(GkSystem), P, P, P]),
    "gk_test_cholesky_f32_split16": (C.c_int, [P, P, C.c_int64, C.c_int64, P, C.POINTER(C.c_int32), P]),
    "gk_test_syrk_split16": (C.c_int, [P, C.c_int64, C.c_int64, P, P, P, P, P, C.c_int64, C.c_int32, P]),
    "gk_finite_fault_values": (C.c_int, [P, C.c_int64, P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), P, P]),
    "gk_gravity_tz": (C.c_int, [P, C.c_int64, P, P, P, C.c_double, P, P]),
    "gk_gravity": (C.c_int, [P, C.c_int64, C.c_int64, P, P, C.c_int32, C.POINTER(C.c_double), P, P]),
    "gk_block_values": (C.c_int, [P, P, C.c_int64, C.c_int32, C.POINTER(C.c_double), C.c_double, P, P]),
    "gk_combine_stacks": (C.c_int, [P, C.c_int32, P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), P, P,
                                    P]),
    "gk_octree_points": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64), P, C.c_int64, C.c_int32,
                                   P, P]),
    "gk_octree_refine": (C.c_int, [P, C.c_int32, P, C.c_int64, P, C.POINTER(C.c_int64), P, P,
                                   C.POINTER(C.c_int64), P]),
    "gk_test_trsv_set_error": (C.c_int, [P, C.c_int32, P]),
}

_lib = None


def load():
    """Load libgkhip.so and bind every declared symbol; raises if anything is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m gempy_amd.build` "
                           "(the hip backend has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)          # AttributeError if the symbol is missing
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GkError(RuntimeError):
    def __init__(self, fn, code):
        msg = load().gk_last_error().decode(errors="replace")
        super().__init__(f"{fn} failed: {GK_ERR.get(code, code)}: {msg}")
        self.code = code


def check(fn: str, code: int) -> None:
    if code != GK_OK:
        raise GkError(fn, code)
