"""Solver vocabulary and shared machinery, device-resident.

Same names, fields, defaults, messages and decisions as
``ttnumpy/solvers/common.py`` (ttnumpy 0.3.0); the arithmetic runs in libttx:

* ``StopReason`` / ``LocalSolver`` / ``TruncationMode`` / ``SolverInfo``
  (``common.py:36-115``);
* ``relative_residual`` -- ``A x - b`` formed as a train on the device and
  normed by a Q-less right QR sweep (``common.py:118-146``);
* environments (``common.py:149-217``) via :mod:`ttnumpy_amd.device`;
* ``_SolverSettings.resolve`` (``common.py:301-400``);
* the sweep monitor (``common.py:447-681``);
* the local solve: LU on the explicit matrix, or device MINRES with the
  true-residual restart loop (``common.py:729-809``);
* the rank rules (``common.py:812-929``), evaluated on the device.
"""

from __future__ import annotations

import logging
import math
import time
from dataclasses import dataclass, field
from enum import Enum
from typing import List, Optional

import torch

from .. import _lib
from .. import device as D
from ..device import P
from ..tensor_train import DimensionMismatch, TensorTrain, WrongTTType

logger = logging.getLogger("ttnumpy_amd")


class StopReason(str, Enum):
    """Why an iterative TT solve stopped (``common.py:36-53``)."""

    CONVERGED = "converged"
    STAGNATED = "stagnated"
    MAX_SWEEPS = "max_sweeps"

    def __str__(self) -> str:
        return str.__str__(self)


class LocalSolver(str, Enum):
    """Local-system solver selection (``common.py:56-68``)."""

    AUTO = "auto"
    DIRECT = "direct"
    MINRES = "minres"
    GMRES = "gmres"

    def __str__(self) -> str:
        return str.__str__(self)


class TruncationMode(str, Enum):
    """Rule that picks the truncation rank of a solved core (``common.py:71-82``)."""

    RESIDUAL = "residual"
    THRESHOLD = "threshold"

    def __str__(self) -> str:
        return str.__str__(self)


@dataclass
class SolverInfo:
    """Convergence information of a TT solve (``common.py:85-115``)."""

    residuals: List[float] = field(default_factory=list)
    num_sweeps_run: int = 0
    stop_reason: StopReason = StopReason.MAX_SWEEPS
    max_bonds: List[int] = field(default_factory=list)
    local_solves: int = 0
    local_fallbacks: int = 0
    local_inexact: int = 0

    @property
    def converged(self) -> bool:
        return self.stop_reason == StopReason.CONVERGED


def residual_train(A: TensorTrain, b: TensorTrain, x: TensorTrain) -> TensorTrain:
    """``A.dot(x).add(b.multiply(-1))`` built core by core on the device."""
    d = A.n_cores
    k_neg = b.orth_center if b.orth_center is not None else 0
    cores = []
    for i in range(d):
        a, xc, bc = A.cores[i], x.cores[i], b.cores[i]
        Ra, nr, _, Rb = a.shape
        rx, _, _, rx2 = xc.shape
        rb, _, _, rb2 = bc.shape
        L = 1 if i == 0 else Ra * rx + rb
        R = 1 if i == d - 1 else Rb * rx2 + rb2
        c = D.zeros(L, nr, 1, R)
        D.ttm_core(a, xc, P(c), nr * R, R, 1)
        off = (L - rb) * nr * R + (R - rb2)
        D.copy_nd([rb, nr, rb2], P(bc), [nr * rb2, rb2, 1], P(c, off), [nr * R, R, 1],
                  -1.0 if i == k_neg else 1.0)
        cores.append(c)
    return TensorTrain(cores)


def relative_residual_device(A: TensorTrain, b: TensorTrain, x: TensorTrain,
                             b_norm: Optional[torch.Tensor] = None) -> torch.Tensor:
    r = residual_train(A, b, x)
    rn = r.tt_norm_device()
    if b_norm is None:
        b_norm = b.tt_norm_device()
    out = D.empty(1)
    # ||b|| == 0 returns the absolute residual (common.py:144-146)
    D.copy_nd([1], P(rn), [1], P(out), [1])
    bn = float(b_norm.item())
    if bn == 0.0:
        return out
    D.copy_nd([1], P(rn), [1], P(out), [1], 1.0 / bn)
    return out


def relative_residual(A: TensorTrain, b: TensorTrain, x: TensorTrain, b_norm: Optional[float] = None) -> float:
    """``||A x - b|| / ||b||`` in TT arithmetic (``common.py:118-146``)."""
    r = residual_train(A, b, x)
    rn = float(r.tt_norm_device().item())
    if b_norm is None:
        b_norm = b.tt_norm()
    if b_norm == 0.0:
        return rn
    return rn / b_norm


# ---------------------------------------------------------------- environments
def _c(t: TensorTrain, k: int) -> torch.Tensor:
    return t.cores[k]


def prepare_left_environment(ind: int, env: list, frame: TensorTrain, target: TensorTrain,
                             A: Optional[TensorTrain] = None) -> None:
    """``env[ind]`` from ``env[ind-1]`` (``common.py:149-188``); real cores (conj is identity)."""
    if ind == 0:
        env[0] = D.ones_env(A is not None)
        return
    env[ind] = D.env_left(env[ind - 1], _c(frame, ind - 1), _c(target, ind - 1),
                          _c(A, ind - 1) if A is not None else None)


def prepare_right_environment(ind: int, env: list, frame: TensorTrain, target: TensorTrain,
                              A: Optional[TensorTrain] = None) -> None:
    """Mirror (``common.py:191-217``)."""
    if ind == len(frame) - 1:
        env[-1] = D.ones_env(A is not None)
        return
    env[ind] = D.env_right(env[ind + 1], _c(frame, ind + 1), _c(target, ind + 1),
                           _c(A, ind + 1) if A is not None else None)


# ---------------------------------------------------------------- settings
_LOCAL_TOL_FACTOR = 0.1
_DEFAULT_LOCAL_RTOL = 1e-8
_MAX_LOCAL_RESTARTS = 4
_RESTART_RTOL_SHRINK = 1e-2
_DEFAULT_STAGNATION_FACTOR = 0.9


@dataclass(frozen=True)
class _SolverSettings:
    """Validated and derived options (``common.py:244-400``)."""

    max_sweeps: int
    residual_tol: Optional[float]
    stagnation_factor: Optional[float]
    iterative_solver: Optional[LocalSolver]
    direct_solve_size: int
    track_residual: bool
    local_rtol: Optional[float]
    allow_dense_fallback: bool = False
    truncation_steps: int = 1
    symmetric: bool = False

    @property
    def truncation_mode(self) -> TruncationMode:
        return TruncationMode.RESIDUAL if self.residual_tol is not None else TruncationMode.THRESHOLD

    @property
    def truncation_tol(self) -> Optional[float]:
        if self.residual_tol is None:
            return None
        return self.residual_tol / math.sqrt(self.truncation_steps)

    def use_iterative(self, size: int) -> bool:
        return self.iterative_solver is not None and size > self.direct_solve_size

    @classmethod
    def resolve(cls, *, max_sweeps, residual_tol, stagnation_factor, local_solver, symmetric,
                direct_solve_size, return_info, local_rtol=None, allow_dense_fallback=False,
                truncation_steps=1) -> "_SolverSettings":
        if residual_tol is not None and residual_tol <= 0.0:
            raise ValueError(f"residual_tol must be positive, got {residual_tol}")
        if truncation_steps < 1:
            raise ValueError(f"truncation_steps must be positive, got {truncation_steps}")
        if isinstance(stagnation_factor, str):
            if stagnation_factor != "auto":
                raise ValueError(
                    "stagnation_factor must be a float in (0, 1], None or "
                    f"'auto', got {stagnation_factor!r}")
            stagnation_factor = (_DEFAULT_STAGNATION_FACTOR
                                 if residual_tol is not None or return_info else None)
        if stagnation_factor is not None and not 0.0 < stagnation_factor <= 1.0:
            raise ValueError(f"stagnation_factor must be in (0, 1], got {stagnation_factor}")
        track = (return_info or residual_tol is not None
                 or (stagnation_factor is not None and max_sweeps > 1))
        if local_rtol is not None and local_rtol <= 0.0:
            raise ValueError(f"local_rtol must be positive, got {local_rtol}")
        try:
            local_solver = LocalSolver(local_solver)
        except ValueError as exc:
            allowed = ", ".join(repr(m.value) for m in LocalSolver)
            raise ValueError(f"local_solver must be one of {allowed}, got {local_solver!r}") from exc
        if local_solver is LocalSolver.AUTO:
            it = LocalSolver.MINRES if symmetric else LocalSolver.GMRES
        elif local_solver is LocalSolver.DIRECT:
            it = None
        else:
            it = local_solver
        if it is None:
            lr = None
        elif local_rtol is not None:
            lr = local_rtol
        elif residual_tol is not None:
            lr = _LOCAL_TOL_FACTOR * residual_tol
        else:
            lr = _DEFAULT_LOCAL_RTOL
        if lr is not None and residual_tol is not None and lr > 0.5 * residual_tol:
            logger.warning("local_rtol %.2e exceeds half of residual_tol %.2e: the local solves "
                           "may be too loose to drive the truncation budget", lr, residual_tol)
        return cls(max_sweeps=max_sweeps, residual_tol=residual_tol,
                   stagnation_factor=stagnation_factor, iterative_solver=it,
                   direct_solve_size=direct_solve_size, track_residual=track, local_rtol=lr,
                   allow_dense_fallback=allow_dense_fallback, truncation_steps=truncation_steps,
                   symmetric=bool(symmetric))


@dataclass
class _SweepStats:
    """Per-sweep tallies (``common.py:403-444``)."""

    max_rank: Optional[int] = None
    rank_decisions: int = 0
    rank_capped: int = 0
    local_solves: int = 0
    local_fallbacks: int = 0
    local_inexact: int = 0

    def record_rank(self, capped: bool) -> None:
        self.rank_decisions += 1
        self.rank_capped += int(capped)

    def record_local_solve(self, fell_back: bool = False, inexact: bool = False) -> None:
        self.local_solves += 1
        self.local_fallbacks += int(fell_back)
        self.local_inexact += int(inexact)

    def clear(self) -> None:
        self.rank_decisions = self.rank_capped = 0
        self.local_solves = self.local_fallbacks = self.local_inexact = 0


class _SweepMonitor:
    """Per-sweep residual, stop test and counters (``common.py:447-681``)."""

    def __init__(self, settings: _SolverSettings, name: str, A: TensorTrain, b: TensorTrain):
        self.settings = settings
        self.name = name
        self.A, self.b = A, b
        self.b_norm = b.tt_norm() if settings.track_residual else None
        self.info = SolverInfo()
        self._t = time.perf_counter()

    def record(self, sweep: int, x: TensorTrain, max_bond: Optional[int] = None,
               stats: Optional[_SweepStats] = None) -> bool:
        info = self.info
        info.num_sweeps_run = sweep + 1
        if max_bond is not None:
            info.max_bonds.append(max_bond)
        if stats is not None:
            info.local_solves += stats.local_solves
            info.local_fallbacks += stats.local_fallbacks
            info.local_inexact += stats.local_inexact
            stats.clear()
        try:
            if not self.settings.track_residual:
                return False
            res = relative_residual(self.A, self.b, x, b_norm=self.b_norm)
            info.residuals.append(res)
            logger.debug("%s sweep %d/%d: relative residual %.3e, max bond %d, %.3fs", self.name,
                         sweep + 1, self.settings.max_sweeps, res, max(x.bonds),
                         time.perf_counter() - self._t)
            if self.settings.residual_tol is not None and res <= self.settings.residual_tol:
                info.stop_reason = StopReason.CONVERGED
                return True
            f = self.settings.stagnation_factor
            if f is not None and len(info.residuals) >= 2 and info.residuals[-1] > f * info.residuals[-2]:
                info.stop_reason = StopReason.STAGNATED
                logger.warning("%s stagnated at sweep %d: relative residual %.3e vs %.3e previous",
                               self.name, sweep + 1, info.residuals[-1], info.residuals[-2])
                return True
            return False
        finally:
            self._t = time.perf_counter()

    def finish(self) -> None:
        if not self.info.residuals:
            return
        if self.settings.residual_tol is not None and self.info.stop_reason is StopReason.MAX_SWEEPS:
            logger.warning("%s did not reach residual_tol %.2e after %d sweep(s) (final relative "
                           "residual %.3e)", self.name, self.settings.residual_tol,
                           self.info.num_sweeps_run, self.info.residuals[-1])


def _validate_solver_inputs(A: TensorTrain, b: TensorTrain, x0: TensorTrain) -> None:
    """Structural checks (``common.py:684-726``)."""
    if x0.is_ttm():
        raise WrongTTType("Initial guess x0 must be a TensorTrain, not a TTM.")
    if b.is_ttm():
        raise WrongTTType("Right-hand side b must be a TensorTrain, not a TTM.")
    if not A.is_ttm():
        raise WrongTTType("Matrix A must be a TTM, not a TensorTrain.")
    if not len(A) == len(b) == len(x0):
        raise DimensionMismatch(
            f"A, b and x0 must have the same number of cores, got {len(A)}, {len(b)} and {len(x0)}")
    if A.row_modes != A.col_modes:
        raise DimensionMismatch(
            f"Solver requires a square operator, got row modes {A.row_modes} and column modes {A.col_modes}")
    if A.row_modes != b.row_modes:
        raise DimensionMismatch(
            f"Row modes of A must match physical dimensions of b: {A.row_modes} != {b.row_modes}")
    if A.col_modes != x0.row_modes:
        raise DimensionMismatch(
            f"Column modes of A must match physical dimensions of x0: {A.col_modes} != {x0.row_modes}")


# ---------------------------------------------------------------- local solves
# MINRES steps queued between polls of the device stop flag: _POLL_FIRST, then
# doubling up to _POLL (most local solves stop within a few steps, and the
# steps queued past the stop still cost their launches)
_POLL_FIRST = 2
_POLL = 32


def minres_device(op, b: torch.Tensor, x0: Optional[torch.Tensor], rtol: float,
                  state_out: Optional[list] = None) -> torch.Tensor:
    """``scipy.sparse.linalg.minres(op, b, x0=x0, rtol=rtol)`` on the device.

    The recurrence and stopping tests run in ``ttx_minres_step_f64``; the
    operator is applied between steps; the host only polls the stop flag.
    """
    n = op.rows
    vec = D.empty(8 * n)
    st = D.empty(_lib.TTX_MINRES_STATE)
    yA = D.empty(n)
    L = D.lib()
    s = D.stream()
    ax0 = None
    if x0 is not None:
        op.apply(P(x0), 1, P(yA))
        ax0 = P(yA)
    _lib.check(L.ttx_minres_init_f64(s, n, P(b), P(x0) if x0 is not None else None, ax0, rtol,
                                     5 * n, P(vec), P(st)), "ttx_minres_init_f64")
    maxiter = 5 * n
    done = 0
    chunked = hasattr(op, "minres_steps")
    poll = _POLL_FIRST
    while True:
        istop, itn = st[:2].tolist()
        if istop != 0.0:
            break
        chunk = min(poll, maxiter - done)
        poll = min(2 * poll, _POLL)
        if chunk <= 0:
            break
        if chunked:
            op.minres_steps(P(yA), P(vec), P(st), chunk)
        else:
            for _ in range(chunk):
                op.apply(P(vec, n), 1, P(yA))
                _lib.check(L.ttx_minres_step_f64(s, n, P(yA), P(vec), P(st)), "ttx_minres_step_f64")
        done += chunk
    if chunked:     # the matvecs queued after the stop test returned at once
        D.add_flops(itn * op.flops_per_apply(1))
    if state_out is not None:
        state_out.append(st.cpu().numpy())
    return vec[:n].clone()


_GMRES_RESTART = 20   # scipy.sparse.linalg.gmres's default restart length


def gmres_device(op, b: torch.Tensor, x0: Optional[torch.Tensor], rtol: float,
                 state_out: Optional[list] = None) -> torch.Tensor:
    """``scipy.sparse.linalg.gmres(op, b, x0=x0, rtol=rtol, atol=0.0)`` on the device.

    The Arnoldi steps, rotations, cycle updates and stopping tests run in the
    ``ttx_gmres_*`` kernels; the operator is applied between them (for the
    one-site operator a whole cycle is issued by one
    ``ttx_gmres_local_cycle_f64`` call); the host polls the done flag once
    per cycle."""
    n = op.rows
    m = min(_GMRES_RESTART, n)
    maxiter = 10 * n
    L = D.lib()
    s = D.stream()
    work = D.empty(int(L.ttx_gmres_work_size(n, m)))
    st = D.empty(_lib.TTX_GMRES_STATE)
    xoff, axoff = (m + 1) * n, (m + 2) * n
    ax0 = None
    if x0 is not None:
        op.apply(P(x0), 1, P(work, axoff))
        ax0 = P(work, axoff)
    _lib.check(L.ttx_gmres_init_f64(s, n, m, P(b), P(x0) if x0 is not None else None, ax0, rtol,
                                    maxiter, P(work), P(st)), "ttx_gmres_init_f64")
    chunked = hasattr(op, "gmres_cycle")
    cycles = 0
    while cycles < maxiter and float(st[0].item()) == 0.0:
        if chunked:
            op.gmres_cycle(m, P(b), P(work), P(st))
        else:
            for col in range(m):
                op.apply(P(work, col * n), 1, P(work, (col + 1) * n))
                _lib.check(L.ttx_gmres_arnoldi_f64(s, n, m, col, P(work), P(st)),
                           "ttx_gmres_arnoldi_f64")
            _lib.check(L.ttx_gmres_finish_f64(s, n, m, P(work), P(st)), "ttx_gmres_finish_f64")
            op.apply(P(work, xoff), 1, P(work, axoff))
            _lib.check(L.ttx_gmres_check_f64(s, n, m, P(b), P(work), P(st)), "ttx_gmres_check_f64")
        cycles += 1
    if chunked:     # the matvecs queued after an inner stop returned at once
        inner, done_cycles = st[13].item(), st[11].item()
        D.add_flops((inner + done_cycles) * op.flops_per_apply(1))
    if state_out is not None:
        state_out.append(st.cpu().numpy())
    return work[xoff:xoff + n].clone()


def solve_local_system(op, b_loc: torch.Tensor, *, iterative_solver, local_rtol, warm_start=None,
                       explicit_fallback=None, stats: Optional[_SweepStats] = None,
                       spd_first: bool = False):
    """Restarted Krylov solve with the true-residual check (``common.py:776-809``).

    Returns ``(u, op)``.  The direct path is handled by the caller (LU on the
    explicit matrix), as the reference does at ``common.py:766-774``; with
    ``explicit_fallback`` true a solve that misses its target after the
    restarts is redone that way (``common.py:794-809``).
    """
    n = op.rows
    tgt = D.norm(b_loc)
    target = local_rtol * float(tgt.item())
    u = warm_start
    rt = local_rtol
    res = D.empty(1)
    au = D.empty(n)
    for _ in range(_MAX_LOCAL_RESTARTS):
        if iterative_solver is LocalSolver.GMRES:
            u = gmres_device(op, b_loc, u, rt)
        else:
            u = minres_device(op, b_loc, u, rt)
        op.apply(P(u), 1, P(au))
        D.diff_nrm2(n, P(au), P(b_loc), P(res))
        if float(res.item()) <= target:
            if stats is not None:
                stats.record_local_solve(fell_back=False)
            return u, op
        rt = max(rt * _RESTART_RTOL_SHRINK, 1e-15)
    if not explicit_fallback:
        if stats is not None:
            stats.record_local_solve(inexact=True)
        return u, op
    if stats is not None:
        stats.record_local_solve(fell_back=True)
    u = direct_solve(op, b_loc, spd_first=spd_first)
    return u, op


def direct_solve(op, b_loc: torch.Tensor, spd_first: bool = False) -> torch.Tensor:
    """Dense solve of [M | b], M the explicit matrix of ``op`` (``common.py:766-774``).

    LU with partial pivoting; with ``spd_first`` (symmetric problems) a
    Cholesky factorization is tried first and LU runs only when M turns out
    not to be positive definite.  Both give the solution scipy.linalg.solve
    gives, up to rounding.  The factorization status is left on
    ``op.lu_info`` (device int32) for the caller's next synchronisation.
    """
    n = op.rows
    Mb = D.empty(n * (n + 1))
    u = D.empty(n)
    info = torch.zeros(1, dtype=torch.int32, device=D.device())

    def assemble():
        op.explicit(Mb, n + 1)
        D.copy_nd([n], P(b_loc), [1], P(Mb, n), [n + 1])

    assemble()
    if spd_first:
        D.chol_solve(Mb, n, u, info)
        if int(info.item()) != _lib.TTX_ENOTSPD:
            op.lu_info = info
            return u
        assemble()
        info.zero_()
    D.lu_solve(Mb, n, u, info)
    op.lu_info = info
    return u
