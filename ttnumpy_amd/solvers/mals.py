"""Two-site MALS on the device (``ttnumpy/solvers/mals.py:61-292``,
``ttnumpy/solvers/two_site.py:28-241``).

A two-site system over cores (k, k+1) is the one-site system of the merged
operator core A_k A_{k+1} (rows (bond, m_k, m_{k+1}, bond), columns
(bond, n_k, n_{k+1}, bond)) with the merged right-hand-side core b_k b_{k+1}:
the local matrix, the matrix-free apply, the direct and MINRES solves and the
rank rules are the one-site device routines on the merged cores.
"""

from __future__ import annotations

from typing import Optional, Tuple, Union

import torch

from .. import device as D
from ..device import P
from ..tensor_train import TensorTrain, ones
from .amen import _choose_rank, _Decision, _read_decision
from .common import (LocalSolver, SolverInfo, TruncationMode, _SolverSettings, _SweepMonitor,
                     _SweepStats, _validate_solver_inputs, direct_solve, prepare_left_environment,
                     prepare_right_environment, solve_local_system)


def _merge_pair(c0: torch.Tensor, c1: torch.Tensor) -> torch.Tensor:
    """(r, m0, n0, s) and (s, m1, n1, t) as one core (r, m0 m1, n0 n1, t)."""
    r, m0, n0, s = c0.shape
    _, m1, n1, t = c1.shape
    both = D.empty(r * m0 * n0, m1 * n1 * t)        # (r, m0, n0, m1, n1, t)
    D.gemm(1, r * m0 * n0, m1 * n1 * t, s, 1.0, P(c0), s, 1, 0, P(c1), m1 * n1 * t, 1, 0, 0.0,
           P(both), m1 * n1 * t, 1, 0)
    out = D.empty(r, m0 * m1, n0 * n1, t)
    # (r, m0, n0, m1, n1, t) -> (r, m0, m1, n0, n1, t)
    D.copy_nd([r, m0, n0, m1, n1, t], P(both),
              [m0 * n0 * m1 * n1 * t, n0 * m1 * n1 * t, m1 * n1 * t, n1 * t, t, 1], P(out),
              [m0 * m1 * n0 * n1 * t, m1 * n0 * n1 * t, n1 * t, n0 * n1 * t, t, 1])
    return out


def _supercore(left: torch.Tensor, right: torch.Tensor) -> torch.Tensor:
    """(r, n0, s) x (s, n1, t) -> flat (r, n0, n1, t)."""
    r, n0, s = left.shape[0], left.shape[1], left.shape[-1]
    n1, t = right.shape[1], right.shape[-1]
    out = D.empty(r * n0 * n1 * t)
    D.gemm(1, r * n0, n1 * t, s, 1.0, P(left.contiguous()), s, 1, 0, P(right.contiguous()), n1 * t,
           1, 0, 0.0, P(out), n1 * t, 1, 0)
    return out


def _solve_local_two_site(k: int, x: TensorTrain, A: TensorTrain, b: TensorTrain, lA: list,
                          rA: list, lb: list, rb: list, settings: _SolverSettings, *,
                          error_threshold: float, max_rank: Optional[int], residual_damp: float,
                          center=None, stats: Optional[_SweepStats] = None):
    """Assemble, solve and SVD-split the two-site system (``two_site.py:147-241``).

    Returns ``(U, S, Vt, K, rows, cols, rank)``; ``U`` is rows x K, ``Vt`` K x cols.
    """
    rows = x.bonds[k] * x.row_modes[k]
    cols = x.row_modes[k + 1] * x.bonds[k + 2]
    size = rows * cols
    Ak = _merge_pair(A.cores[k], A.cores[k + 1])
    bk = _merge_pair(b.cores[k], b.cores[k + 1])
    fl, fr = lb[k].shape[1], rb[k + 1].shape[1]
    b_loc = D.empty(fl * bk.shape[1] * fr)
    D.local_rhs(lb[k], bk, rb[k + 1], P(b_loc))
    op = D.LocalOp(lA[k], Ak, rA[k + 1])
    mode = settings.truncation_mode
    if settings.use_iterative(size):
        left = x.cores[k].reshape(x.cores[k].shape[0], x.cores[k].shape[1], -1)
        right = x.cores[k + 1].reshape(x.cores[k + 1].shape[0], x.cores[k + 1].shape[1], -1)
        if center is not None:
            if center[0] == k:
                left = center[1]
            elif center[0] == k + 1:
                right = center[1]
        warm = None
        if left.shape[-1] == right.shape[0]:
            warm = _supercore(left, right)
            if warm.numel() != size:
                warm = None
        u, op = solve_local_system(op, b_loc, iterative_solver=settings.iterative_solver,
                                   local_rtol=settings.local_rtol, warm_start=warm,
                                   explicit_fallback=settings.allow_dense_fallback, stats=stats,
                                   spd_first=settings.symmetric)
    else:
        u = direct_solve(op, b_loc, spd_first=settings.symmetric)
    dec = _Decision()
    exact = D.empty(1)
    if mode is TruncationMode.RESIDUAL:
        tmp = D.empty(size)
        op.apply(P(u), 1, P(tmp))
        D.diff_nrm2(size, P(tmp), P(b_loc), P(exact))
    U, S, Vt = D.svd(rows, cols, P(u), cols, 1, dec.buf, 0)
    K = min(rows, cols)
    _choose_rank(mode, op, b_loc, U, S, Vt, rows, cols, K, exact, error_threshold, max_rank,
                 settings.truncation_tol, residual_damp, dec)
    rank, capped = _read_decision(dec, op, mode, max_rank, K)
    if stats is not None:
        stats.record_rank(capped)
    return U, S, Vt, K, rows, cols, rank


def mals(A: TensorTrain, b: TensorTrain, x0: Optional[TensorTrain] = None, inplace: bool = False,
         max_rank: Optional[int] = None, max_sweeps: int = 2, residual_tol: Optional[float] = None,
         residual_damp: float = 2.0, stagnation_factor: Union[float, str, None] = "auto",
         error_threshold: float = 1e-7, local_solver: Union[str, LocalSolver] = "auto",
         direct_solve_size: int = 1000, local_rtol: Optional[float] = None,
         allow_dense_fallback: bool = False, symmetric: bool = False, return_info: bool = False
         ) -> Union[TensorTrain, Tuple[TensorTrain, SolverInfo]]:
    """Modified ALS (two-site) for ``A x = b`` (``mals.py:61-292``).

    Left to right over the pairs (k, k+1), k < d-2: core k <- U_r and the
    block S_r V_r is kept as the centre; right to left over all pairs: core
    k+1 <- V_r, the block U_r S_r kept, and at k = 0 stored as core 0.  The
    residual budget is shared by the 2d - 3 truncations of a sweep.
    """
    if max_rank is not None and max_rank <= 0:
        raise ValueError(f"max_rank must be a positive integer, got {max_rank}")
    if residual_damp < 1.0:
        raise ValueError(f"residual_damp must be >= 1, got {residual_damp}")
    settings = _SolverSettings.resolve(
        max_sweeps=max_sweeps, residual_tol=residual_tol, stagnation_factor=stagnation_factor,
        local_solver=local_solver, symmetric=symmetric, direct_solve_size=direct_solve_size,
        return_info=return_info, local_rtol=local_rtol, allow_dense_fallback=allow_dense_fallback,
        truncation_steps=max(2 * A.n_cores - 3, 1))
    if x0 is None:
        x0 = ones(b.physical_dims(), ranks=b.bonds)
    _validate_solver_inputs(A, b, x0)
    x = x0 if inplace else x0.copy()
    if x.orth_center != 0:
        x.orthonormalize(canonical="right", inplace=True)
    d = len(x)
    rA, rb, lA, lb = [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, 0, -1):
        prepare_right_environment(i, rA, x, x, A)
        prepare_right_environment(i, rb, x, b)
    monitor = _SweepMonitor(settings, "MALS", A, b)
    stats = _SweepStats(max_rank=max_rank)
    center = None
    kw = dict(error_threshold=error_threshold, max_rank=max_rank, residual_damp=residual_damp)
    for sweep in range(settings.max_sweeps):
        for k in range(d - 1):
            prepare_left_environment(k, lA, x, x, A)
            prepare_left_environment(k, lb, x, b)
            if k == d - 2:
                continue
            U, S, Vt, K, rows, cols, nb = _solve_local_two_site(
                k, x, A, b, lA, rA, lb, rb, settings, center=center, stats=stats, **kw)
            n0, n1 = x.row_modes[k], x.row_modes[k + 1]
            sv = Vt[:nb].clone()                                     # S_r V_r
            D.scale_cols(cols, nb, P(sv), 1, cols, P(S))
            center = (k + 1, sv.reshape(nb, n1, x.bonds[k + 2]))
            core = D.empty(x.bonds[k], n0, 1, nb)
            D.copy_nd([rows, nb], P(U), [K, 1], P(core), [nb, 1])
            x.cores[k] = core
            x.bonds[k + 1] = nb
            x.orth_center = k + 1
        for k in range(d - 2, -1, -1):
            prepare_right_environment(k + 1, rA, x, x, A)
            prepare_right_environment(k + 1, rb, x, b)
            U, S, Vt, K, rows, cols, nb = _solve_local_two_site(
                k, x, A, b, lA, rA, lb, rb, settings, center=center, stats=stats, **kw)
            n0, n1 = x.row_modes[k], x.row_modes[k + 1]
            us = D.empty(rows, nb)                                   # U_r S_r
            D.copy_nd([rows, nb], P(U), [K, 1], P(us), [nb, 1])
            D.scale_cols(rows, nb, P(us), nb, 1, P(S))
            center = (k, us.reshape(x.bonds[k], n0, nb))
            x.cores[k + 1] = Vt[:nb].reshape(nb, n1, 1, x.bonds[k + 2]).clone()
            x.bonds[k + 1] = nb
            if k == 0:
                x.cores[0] = us.reshape(x.bonds[0], n0, 1, nb)
            x.orth_center = k
        if monitor.record(sweep, x, max_bond=max(x.bonds), stats=stats):
            break
    x.max_rank = max(x.bonds)
    monitor.finish()
    if return_info:
        return x, monitor.info
    return x
