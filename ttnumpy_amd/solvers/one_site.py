"""One-site local algebra on the device (``ttnumpy/solvers/one_site.py:27-209``)."""

from __future__ import annotations

from typing import Optional

import torch

from .. import device as D
from ..device import P
from ..tensor_train import TensorTrain
from .common import _SolverSettings, _SweepStats, direct_solve, solve_local_system


def local_rhs_one_site(ind: int, left_env_b: list, right_env_b: list, b: TensorTrain) -> torch.Tensor:
    """Flat ``b_loc`` indexed (row_bond, mode, right_bond) (``one_site.py:27-46``)."""
    fl = left_env_b[ind].shape[1]
    fr = right_env_b[ind].shape[1]
    n = b.cores[ind].shape[1]
    out = D.empty(fl * n * fr)
    D.local_rhs(left_env_b[ind], b.cores[ind], right_env_b[ind], P(out))
    return out


def local_explicit_matrix_one_site(ind: int, left_env_A: list, right_env_A: list,
                                   A: TensorTrain) -> D.ExplicitOp:
    """Explicit local matrix, column-major (``one_site.py:49-75``)."""
    op = D.LocalOp(left_env_A[ind], A.cores[ind], right_env_A[ind])
    n = op.rows
    M = D.empty(n * n)
    op.explicit(M, n)
    return D.ExplicitOp(M, n)


def _solve_local_one_site(k: int, x: TensorTrain, A: TensorTrain, b: TensorTrain, left_env_A: list,
                          right_env_A: list, left_env_b: list, right_env_b: list,
                          settings: _SolverSettings, preserve_operator: bool = False,
                          stats: Optional[_SweepStats] = None):
    """Assemble and solve the local system at ``k`` (``one_site.py:118-178``).

    Returns ``(u, operator, b_loc)`` with ``u`` flat (left_bond, mode, right_bond).
    """
    size = x.bonds[k] * x.row_modes[k] * x.bonds[k + 1]
    b_loc = local_rhs_one_site(k, left_env_b, right_env_b, b)
    if settings.use_iterative(size):
        op = D.LocalOp(left_env_A[k], A.cores[k], right_env_A[k])
        warm = x.cores[k].reshape(-1)
        if warm.numel() != size:
            warm = None
        u, op = solve_local_system(op, b_loc, iterative_solver=settings.iterative_solver,
                                   local_rtol=settings.local_rtol, warm_start=warm,
                                   explicit_fallback=settings.allow_dense_fallback, stats=stats,
                                   spd_first=settings.symmetric)
        return u, op, b_loc
    # the explicit matrix goes straight into the LU buffer; the residual and
    # rank tests that follow apply the same operator matrix-free
    op = D.LocalOp(left_env_A[k], A.cores[k], right_env_A[k])
    u = direct_solve(op, b_loc, spd_first=settings.symmetric)
    return u, op, b_loc


def _shift_center_right(x: TensorTrain, k: int, core: torch.Tensor) -> None:
    """QR of (r_l n x r_r); Q at k, R into k+1 (``one_site.py:181-193``)."""
    rl, n, rr = core.shape[0], core.shape[1], core.shape[-1]
    m = rl * n
    kk = min(m, rr)
    Q = D.empty(rl, n, 1, kk)
    R = D.empty(kk, rr)
    D.qr(m, rr, P(core), rr, 1, P(Q), kk, 1, P(R), rr, 1)
    x.cores[k] = Q
    x.cores[k + 1] = D.absorb_left(P(R), kk, rr, x.cores[k + 1])
    x.bonds[k + 1] = kk
    x.orth_center = k + 1


def _shift_center_left(x: TensorTrain, k: int, core: torch.Tensor) -> None:
    """QR of the transposed right unfolding; R^T into k-1 (``one_site.py:196-209``)."""
    rl, n, rr = core.shape[0], core.shape[1], core.shape[-1]
    w = n * rr
    kk = min(w, rl)
    Qt = D.empty(kk, n, 1, rr)
    R = D.empty(kk, rl)
    D.qr(w, rl, P(core), 1, w, P(Qt), 1, w, P(R), rl, 1)
    x.cores[k] = Qt
    x.cores[k - 1] = D.absorb_right(x.cores[k - 1], P(R), rl, kk, m_rs=1, m_cs=rl)
    x.bonds[k] = kk
    x.orth_center = k - 1
