"""Fixed-rank one-site ALS on the device (``ttnumpy/solvers/als.py:56-212``)."""

from __future__ import annotations

from typing import Optional, Tuple, Union

from ..tensor_train import TensorTrain, ones
from .common import (LocalSolver, SolverInfo, _SolverSettings, _SweepMonitor, _SweepStats,
                     _validate_solver_inputs, prepare_left_environment, prepare_right_environment)
from .one_site import _shift_center_left, _shift_center_right, _solve_local_one_site


def als(A: TensorTrain, b: TensorTrain, x0: Optional[TensorTrain] = None, inplace: bool = False,
        max_sweeps: int = 4, residual_tol: Optional[float] = None,
        stagnation_factor: Union[float, str, None] = "auto",
        local_solver: Union[str, LocalSolver] = "auto", direct_solve_size: int = 1000,
        local_rtol: Optional[float] = None, allow_dense_fallback: bool = False,
        symmetric: bool = False, return_info: bool = False
        ) -> Union[TensorTrain, Tuple[TensorTrain, SolverInfo]]:
    """Alternating least squares for ``A x = b`` at the ranks of ``x0`` (``als.py:56-212``)."""
    settings = _SolverSettings.resolve(
        max_sweeps=max_sweeps, residual_tol=residual_tol, stagnation_factor=stagnation_factor,
        local_solver=local_solver, symmetric=symmetric, direct_solve_size=direct_solve_size,
        return_info=return_info, local_rtol=local_rtol, allow_dense_fallback=allow_dense_fallback)
    if x0 is None:
        x0 = ones(b.physical_dims(), ranks=b.bonds)
    _validate_solver_inputs(A, b, x0)
    x = x0 if inplace else x0.copy()
    if x.orth_center != 0:
        x.orthonormalize(canonical="right", inplace=True)
    d = len(x)
    rA, rb, lA, lb = [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, -1, -1):
        prepare_right_environment(i, rA, x, x, A)
        prepare_right_environment(i, rb, x, b)
    monitor = _SweepMonitor(settings, "ALS", A, b)
    stats = _SweepStats()
    for sweep in range(settings.max_sweeps):
        for k in range(d - 1):
            prepare_left_environment(k, lA, x, x, A)
            prepare_left_environment(k, lb, x, b)
            u, _, _ = _solve_local_one_site(k, x, A, b, lA, rA, lb, rb, settings, stats=stats)
            _shift_center_right(x, k, u.reshape(x.bonds[k], x.cores[k].shape[1], x.bonds[k + 1]))
        prepare_left_environment(d - 1, lA, x, x, A)
        prepare_left_environment(d - 1, lb, x, b)
        for k in range(d - 1, -1, -1):
            prepare_right_environment(k, rA, x, x, A)
            prepare_right_environment(k, rb, x, b)
            u, _, _ = _solve_local_one_site(k, x, A, b, lA, rA, lb, rb, settings, stats=stats)
            if k == 0:
                x.cores[0] = u.reshape(x.bonds[0], x.cores[0].shape[1], 1, x.bonds[1])
                x.orth_center = 0
            else:
                _shift_center_left(x, k, u.reshape(x.bonds[k], x.cores[k].shape[1], x.bonds[k + 1]))
        if monitor.record(sweep, x, stats=stats):
            break
    x.max_rank = max(x.bonds)
    monitor.finish()
    if return_info:
        return x, monitor.info
    return x
