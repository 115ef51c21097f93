"""ctypes binding of ``libttx.so`` (declared in ``include/ttx.h`` and ``include/ttx_engine.h``).

The library is the only compute path: if it is missing, or no HIP device is
present, every operation raises instead of falling back to the CPU.
"""

from __future__ import annotations

import ctypes
import os
import warnings

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libttx.so")

TTX_OK = 0
TTX_EINVAL = 1
TTX_ENONFINITE = 2
TTX_ENOCONV = 3
TTX_EHIP = 4
TTX_ESINGULAR = 5
TTX_ENOTSPD = 6
TTX_ESCRATCH = 7
TTX_MINRES_STATE = 32
TTX_GMRES_STATE = 16

_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_p = ctypes.c_void_p
_i32 = ctypes.c_int

# name -> (restype, argtypes)
_SIGNATURES = {
    "ttx_version": (_i32, []),
    "ttx_gemm_f64": (_i32, [_p, _i64, _i64, _i64, _i64, _f64, _p, _i64, _i64, _i64, _p, _i64,
                            _i64, _i64, _f64, _p, _i64, _i64, _i64]),
    "ttx_gemm_f32acc": (_i32, [_p, _i64, _i64, _i64, _i64, _f64, _p, _i64, _i64, _i64, _p, _i64,
                            _i64, _i64, _f64, _p, _i64, _i64, _i64, _p, _p]),
    "ttx_copy_f64": (_i32, [_p, _i32, _p, _p, _p, _p, _p, _f64]),
    "ttx_fill_f64": (_i32, [_p, _p, _i64, _f64]),
    "ttx_axpby_f64": (_i32, [_p, _i64, _f64, _p, _f64, _p, _p]),
    "ttx_nrm2_f64": (_i32, [_p, _i64, _p, _p]),
    "ttx_diff_nrm2_f64": (_i32, [_p, _i64, _p, _p, _p]),
    "ttx_qr_work_size": (_i64, [_i64, _i64]),
    "ttx_qr_f64": (_i32, [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _p]),
    "ttx_svd_work_size": (_i64, [_i64, _i64]),
    "ttx_svd_f64": (_i32, [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _i64,
                           _p, _p]),
    "ttx_svd_noalloc_max_work": (_i64, []),
    "ttx_svd_f64_noalloc": (_i32, [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64,
                                   _i64, _p]),
    "ttx_lu_work_size": (_i64, [_i64]),
    "ttx_lu_solve_f64": (_i32, [_p, _i64, _p, _i64, _p, _p, _p]),
    "ttx_chol_solve_f64": (_i32, [_p, _i64, _p, _i64, _p, _p]),
    "ttx_minres_init_f64": (_i32, [_p, _i64, _p, _p, _p, _f64, _i64, _p, _p]),
    "ttx_minres_step_f64": (_i32, [_p, _i64, _p, _p, _p]),
    "ttx_local_apply_f64": (_i32, [_p, _p, _i64, _i64, _i64, _p, _i64, _i64, _i64, _p, _i64, _i64,
                                   _p, _i64, _f64, _f64, _p, _p, _p]),
    "ttx_minres_local_steps_f64": (_i32, [_p, _p, _i64, _i64, _i64, _p, _i64, _i64, _i64, _p, _i64,
                                          _i64, _p, _p, _p, _p, _p, _i64]),
    "ttx_gmres_work_size": (_i64, [_i64, _i64]),
    "ttx_gmres_init_f64": (_i32, [_p, _i64, _i64, _p, _p, _p, _f64, _i64, _p, _p]),
    "ttx_gmres_arnoldi_f64": (_i32, [_p, _i64, _i64, _i64, _p, _p]),
    "ttx_gmres_finish_f64": (_i32, [_p, _i64, _i64, _p, _p]),
    "ttx_gmres_check_f64": (_i32, [_p, _i64, _i64, _p, _p, _p]),
    "ttx_gmres_local_cycle_f64": (_i32, [_p, _p, _i64, _i64, _i64, _p, _i64, _i64, _i64, _p, _i64,
                                         _i64, _p, _p, _i64, _p, _p, _p]),
    "ttx_truncated_bond_f64": (_i32, [_p, _p, _i64, _f64, _i64, _p]),
    "ttx_truncated_bond_dev_f64": (_i32, [_p, _p, _i64, _p, _f64, _i64, _p]),
    "ttx_rank_by_residual_f64": (_i32, [_p, _i64, _i64, _p, _p, _p, _p, _f64, _f64, _p, _p]),
    "ttx_svd_terms_f64": (_i32, [_p, _i64, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _i64, _p]),
    "ttx_local_matrix_f64": (_i32, [_p, _p, _i64, _i64, _i64, _p, _i64, _i64, _i64, _p, _i64,
                                    _i64, _p, _i64, _i64]),
    "ttx_ttm_core_f64": (_i32, [_p, _p, _i64, _i64, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64,
                                _i64]),
    "ttx_scale_cols_f64": (_i32, [_p, _i64, _i64, _p, _i64, _i64, _p]),
    "ttx_scale_maxabs_f64": (_i32, [_p, _i64, _p]),
    "ttx_hadamard_core_f64": (_i32, [_p, _p, _i64, _i64, _i64, _i64, _p, _i64, _i64, _p]),
    "ttx_ttm_ttm_core_f64": (_i32, [_p, _p, _i64, _i64, _i64, _i64, _p, _i64, _i64, _i64, _p]),
    "ttx_gemm_set_small_tile_limit": (_i32, [_i64]),
    "ttx_gemm_small_tile_limit": (_i64, []),
    "ttx_rank_by_threshold_f64": (_i32, [_p, _p, _i64, _f64, _i64, _p]),
    "ttx_engine_slot_bytes": (_i64, []),
    "ttx_engine_field_offset": (_i64, [_i32]),
    "ttx_engine_plan": (_i32, [_p]),
    "ttx_engine_tick": (_i32, [_p, _p, _i64, _i64, ctypes.c_uint32]),
    "ttx_engine_run": (_i32, [_p, _p, _i64, _i64, _i64]),
    "ttx_engine_profile": (_i32, [_p, _p, _i64, _i64, ctypes.POINTER(_f64)]),
    "ttx_engine_graph_create": (_i32, [_p, _i64, _i64, ctypes.c_uint32, ctypes.POINTER(_p)]),
    "ttx_engine_graph_launch": (_i32, [_p, _p]),
    "ttx_engine_graph_destroy": (_i32, [_p]),
}

_lib = None


class TTXError(RuntimeError):
    """A HIP launch or runtime failure inside libttx."""


def load() -> ctypes.CDLL:
    """Load libttx.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m ttnumpy_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGNATURES)


def check(rc: int, what: str) -> None:
    """Map a ttx status to the exception the reference raises for it."""
    if rc == TTX_OK:
        return
    if rc == TTX_EINVAL:
        raise ValueError(f"{what}: invalid argument")
    if rc == TTX_ESCRATCH:
        raise ValueError(f"{what}: the problem needs more work memory than the fixed scratch "
                         "array holds")
    if rc == TTX_ENONFINITE:
        raise ValueError(f"{what}: NaN or infinite entries")
    if rc == TTX_ENOCONV:
        warnings.warn(f"{what}: did not converge", RuntimeWarning)
        return
    raise TTXError(f"{what}: HIP error (status {rc})")
