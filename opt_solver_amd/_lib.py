"""ctypes binding of libsmfgpu.so (include/smfgpu.h).

The library is built in-tree (``opt_solver_amd/libsmfgpu.so``, see
``__graft_entry__.build``). Loading fails loudly when it is missing: there is
no CPU fallback for the numeric phase."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmfgpu.so")

_P = ctypes.c_void_p
_I = ctypes.c_int32
_L = ctypes.c_int64
_D = ctypes.c_double

SMFGPU_OK = 0
PREC_FP64, PREC_FP32, PREC_BF16, PREC_AUTO, PREC_BF16X3 = 0, 1, 2, 3, 4
MEM_HOST, MEM_DEVICE = 0, 1

STATUS = {0: "Success", 1: "NotPositiveDefinite", 2: "Singular", 3: "MaxPivotDelays", 4: "InertiaExceeded"}


class AnalysisDesc(ctypes.Structure):
    _fields_ = [
        ("n", _I), ("nsup", _I),
        ("sn_start", _P), ("sn_end", _P), ("sn_parent", _P),
        ("child_ptr", _P), ("child_idx", _P),
        ("front_row_ptr", _P), ("front_row_idx", _P),
        ("pcol_ptr", _P), ("prow_idx", _P),
        ("n_orig", _L), ("orig_to_slot", _P), ("perm", _P),
    ]


class FactorOptions(ctypes.Structure):
    _fields_ = [
        ("posdef", _I), ("pivot_u", _D), ("small_pivot", _D), ("panel", _I), ("precision", _I),
        ("mixed_min_front", _I), ("max_negative_pivots", _I), ("value_mem", _I),
        ("work_limit_on", _I), ("work_limit", _D), ("static_pivot_tol", _D),
        ("fallback_static_pivot_tol", _D), ("static_flags", _I), ("selector_mask", _I), ("selector_tol", _D),
        ("scaled", _I), ("collect", _I), ("scale_mem", _I), ("reserved", _I), ("scale", _P),
    ]


class RefineOptions(ctypes.Structure):
    _fields_ = [
        ("method", _I), ("max_iters", _I), ("max_steps", _I), ("restart", _I), ("eta_target", _D),
        ("omega_target", _D), ("value_mem", _I), ("reserved", _I * 3),
    ]


class RefineInfo(ctypes.Structure):
    _fields_ = [
        ("eta", _D), ("omega", _D), ("eta_start", _D), ("omega_start", _D), ("steps", _I), ("method_used", _I),
        ("reached", _I), ("stalled", _I), ("fallback", _I), ("pad", _I), ("reserved", _D * 1),
    ]


REFINE_CLASSICAL, REFINE_AUTO, REFINE_FGMRES = 0, 1, 2


class FactorInfo(ctypes.Structure):
    _fields_ = [
        ("status", _I), ("num_positive", _I), ("num_negative", _I), ("num_zero", _I),
        ("delayed_pivots", _I), ("two_by_two_pivots", _I), ("max_front", _I), ("fronts_reduced", _I),
        ("actual_flops", _D), ("front_work", _D), ("max_abs", _D), ("factor_seconds", _D),
        ("factor_entries", _L), ("static_pivots", _I), ("pad", _I), ("pivot_growth", _D),
        ("reserved", _D * 2),
    ]


EXPORTED = [
    "smfgpu_error_string", "smfgpu_default_options", "smfgpu_analyse", "smfgpu_analysis_describe",
    "smfgpu_analysis_stats", "smfgpu_analysis_free", "smfgpu_plan_create", "smfgpu_plan_free",
    "smfgpu_plan_stats", "smfgpu_factor", "smfgpu_factor_free", "smfgpu_export_factor", "smfgpu_solve",
    "smfgpu_refine", "smfgpu_set_timing", "smfgpu_kernel_times", "smfgpu_plan_tune",
    "smfgpu_factor_front_stats", "smfgpu_launch_profile", "smfgpu_refine_default_options", "smfgpu_refine_ex",
    "smfgpu_solve_noalloc", "smfgpu_factor_precisions", "smfgpu_selector_set", "smfgpu_selector_pairs",
    "smfgpu_selector_train", "smfgpu_selector_predict", "smfgpu_factor_solve", "smfgpu_factor_stall_stats",
    "smfgpu_equilibration_scale",
]
SELECTOR_FEATURES, SELECTOR_OUTPUTS, SELECTOR_WEIGHTS = 10, 3, 1527
STATIC_AS_ZERO = 1
STATIC_ZERO_STAYS_ZERO = 2
STATIC_IN_ORDER = 4
STATUS_SUCCESS = 0
STATUS_NOT_POSDEF = 1
STATUS_SINGULAR = 2
STATUS_MAX_PIVOT_DELAYS = 3
STATUS_INERTIA_EXCEEDED = 4
TUNE_SOLVE_FUSE_ROWS = 1
TUNE_SYRK_AHEAD = 2
TUNE_PANEL_OUTER = 3
TUNE_CB_SEG = 4
TUNE_LDLT_BIG_MIN = 5
TUNE_LDLT_GROUP = 6
TUNE_LDLT_UPDATE_WG = 7
TUNE_LDLT_DEVICE_FRONTS = 8
TUNE_ASM_CHUNK = 9
TUNE_LDLT_DEVICE_WAVE = 10
TUNE_SOLVE_PERSISTENT = 11
TUNE_PANEL_OUTER_REDUCED = 12
TUNE_LDLT_PLANES = 13
TUNE_LDLT_OUTER_WIDE = 14
TUNE_ASM_WAVE_ROWS = 15

_lib = None


class SmfGpuError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SmfGpuError(f"{LIB_PATH} is not built: run __graft_entry__.build() (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    L.smfgpu_error_string.restype = ctypes.c_char_p
    L.smfgpu_default_options.argtypes = [_P]
    L.smfgpu_analyse.argtypes = [_I, _P, _P, _P, _I, _P]
    L.smfgpu_analysis_describe.argtypes = [_P, _P]
    L.smfgpu_analysis_stats.argtypes = [_P, _P]
    L.smfgpu_analysis_free.argtypes = [_P]
    L.smfgpu_plan_create.argtypes = [_P, _I, _P]
    L.smfgpu_plan_free.argtypes = [_P]
    L.smfgpu_plan_stats.argtypes = [_P, _P]
    L.smfgpu_factor.argtypes = [_P, _P, _L, _P, _P, _P]
    L.smfgpu_factor_free.argtypes = [_P]
    L.smfgpu_export_factor.argtypes = [_P, _P, _P, _P, _P, _P]
    L.smfgpu_solve.argtypes = [_P, _P, _I, _I, _I, _I]
    L.smfgpu_solve_noalloc.argtypes = [_P, _P, _I, _I, _I]
    L.smfgpu_factor_precisions.argtypes = [_P, _P]
    L.smfgpu_selector_set.argtypes = [_P, _P]
    L.smfgpu_selector_pairs.argtypes = [_P, _P, _P, _P, _I]
    L.smfgpu_selector_train.argtypes = [_P, _P, _P, _I, _I, ctypes.c_float, _P, _P]
    L.smfgpu_selector_predict.argtypes = [_P, _P, _I, _P]
    L.smfgpu_factor_solve.argtypes = [_P, _P, _L, _P, _P, _P, _I, _I, _P, _I, _P, _P, _P]
    L.smfgpu_equilibration_scale.argtypes = [_P, _P, _L, _I, _I, _P, _I]
    L.smfgpu_refine.argtypes = [_P, _P, _P, _P, _I, _I, _D, _I, _P]
    L.smfgpu_refine_default_options.argtypes = [_P]
    L.smfgpu_refine_ex.argtypes = [_P, _P, _P, _I, _P, _I, _I, _P, _I, _P]
    L.smfgpu_set_timing.argtypes = [_P, _I]
    L.smfgpu_kernel_times.argtypes = [_P, _P]
    L.smfgpu_plan_tune.argtypes = [_P, _I, _L]
    L.smfgpu_factor_front_stats.argtypes = [_P, _P]
    L.smfgpu_factor_stall_stats.argtypes = [_P, _P]
    L.smfgpu_launch_profile.argtypes = [_P, _P, _L]
    L.smfgpu_launch_profile.restype = _L
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != SMFGPU_OK:
        msg = lib().smfgpu_error_string().decode(errors="replace")
        raise SmfGpuError(f"{what}: error {rc}: {msg}")
