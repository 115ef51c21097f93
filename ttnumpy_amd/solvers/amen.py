"""One-site AMEn on the device (``ttnumpy/solvers/amen.py:80-725``).

Same signature, defaults, validation, sweep order, rank rules and
``SolverInfo`` as the reference ``amen``.  Every contraction, factorization
and norm runs in libttx; the host keeps the sweep bookkeeping and takes one
small device->host copy per rank decision (and per sweep for the residual).
"""

from __future__ import annotations

import logging
from math import prod
from typing import Optional, Tuple, Union

import torch

from .. import _lib
from .. import device as D
from ..device import P
from ..tensor_train import TensorTrain, ones, rand
from .common import (LocalSolver, SolverInfo, TruncationMode, _SolverSettings, _SweepMonitor,
                     _SweepStats, _validate_solver_inputs, prepare_left_environment,
                     prepare_right_environment)
from .one_site import (_shift_center_left, _shift_center_right, _solve_local_one_site,
                       local_rhs_one_site)

_log = logging.getLogger("ttnumpy_amd")


def init_residual_train(b: TensorTrain, kickrank: int, seed: Optional[int] = None) -> TensorTrain:
    """Random rank-``kickrank`` train, right-orthonormalised (``amen.py:80-106``)."""
    if kickrank <= 0:
        raise ValueError(f"kickrank must be a positive integer, got {kickrank}")
    z = rand(b.physical_dims(), ranks=kickrank, seed=seed)
    return z.orthonormalize(canonical="right")


def residual_core(ind, left_env_zb, right_env_zb, left_env_zAx, right_env_zAx, A: TensorTrain,
                  b: TensorTrain, x: TensorTrain, x_core: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Projection of ``b - A x`` onto z's frame at ``ind`` (``amen.py:109-159``), (zl, n, zr)."""
    zl = left_env_zb[ind].shape[1]
    zr = right_env_zb[ind].shape[1]
    n = b.cores[ind].shape[1]
    out = D.empty(zl, n, zr)
    D.local_rhs(left_env_zb[ind], b.cores[ind], right_env_zb[ind], P(out))
    op = D.LocalOp(left_env_zAx[ind], A.cores[ind], right_env_zAx[ind])
    xc = x_core if x_core is not None else x.cores[ind]
    op.apply(P(xc), 1, P(out), alpha=-1.0, beta=1.0)
    return out


def fit_residual_train(A: TensorTrain, b: TensorTrain, x: TensorTrain, z: TensorTrain,
                       sweeps: int = 1) -> TensorTrain:
    """ALS fit of z to ``b - A x`` (``amen.py:162-224``)."""
    d = len(z)
    rzb, rzA, lzb, lzA = [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, -1, -1):
        prepare_right_environment(i, rzb, z, b)
        prepare_right_environment(i, rzA, z, x, A)
    for _ in range(sweeps):
        for k in range(d - 1):
            prepare_left_environment(k, lzb, z, b)
            prepare_left_environment(k, lzA, z, x, A)
            core = residual_core(k, lzb, rzb, lzA, rzA, A, b, x)
            _shift_center_right(z, k, core)
        prepare_left_environment(d - 1, lzb, z, b)
        prepare_left_environment(d - 1, lzA, z, x, A)
        for k in range(d - 1, -1, -1):
            prepare_right_environment(k, rzb, z, b)
            prepare_right_environment(k, rzA, z, x, A)
            core = residual_core(k, lzb, rzb, lzA, rzA, A, b, x)
            if k > 0:
                _shift_center_left(z, k, core)
            else:
                z.cores[0] = core.reshape(z.bonds[0], core.shape[1], 1, z.bonds[1])
                z.orth_center = 0
    z.max_rank = max(z.bonds)
    return z


def _slice_last(t: torch.Tensor, m: int) -> torch.Tensor:
    """Contiguous copy of ``t[..., :m]``."""
    lead = t.numel() // t.shape[-1]
    out = D.empty(tuple(t.shape[:-1]) + (m,))
    D.copy_nd([lead, m], P(t), [t.shape[-1], 1], P(out), [m, 1])
    return out


class _Decision:
    """Device int32 slots of one truncation: [svd_info, rank, flag, lu_info]."""

    def __init__(self):
        self.buf = torch.zeros(4, dtype=torch.int32, device=D.device())


def _choose_rank(mode, op, b_loc, U, S, Vt, p, q, K, exact, error_threshold, max_rank,
                 truncation_tol, residual_damp, dec: _Decision):
    if mode is TruncationMode.RESIDUAL:
        limit = K if max_rank is None else min(K, max_rank)
        N = p * q
        terms = D.empty(N * limit)
        _lib.check(D.lib().ttx_svd_terms_f64(D.stream(), p, q, limit, P(U), K, 1, P(S), P(Vt), q, 1,
                                             P(terms)), "ttx_svd_terms_f64")
        contrib = D.empty(N * limit)
        op.apply(P(terms), limit, P(contrib))
        work = D.empty(N)
        _lib.check(D.lib().ttx_rank_by_residual_f64(D.stream(), N, limit, P(S), P(contrib), P(b_loc),
                                                    P(exact), residual_damp, 0.5 * truncation_tol,
                                                    P(work), dec.buf.data_ptr() + 4),
                   "ttx_rank_by_residual_f64")
    else:
        _lib.check(D.lib().ttx_rank_by_threshold_f64(D.stream(), P(S), K, error_threshold,
                                                     0 if max_rank is None else max_rank,
                                                     dec.buf.data_ptr() + 4),
                   "ttx_rank_by_threshold_f64")


def _read_decision(dec: _Decision, op, mode, max_rank, K):
    lu = getattr(op, "lu_info", None)
    if lu is not None:
        dec.buf[3:4].copy_(lu)  # device-to-device, read back with the decision
    info, rank, flag, lu_info = D.sync_ints(dec.buf)
    if lu_info == _lib.TTX_ESINGULAR:
        raise _lib.TTXError("singular local matrix in the direct solve")
    if info == _lib.TTX_ENONFINITE:
        raise ValueError("local solution has NaN or infinite entries")
    if mode is TruncationMode.RESIDUAL:
        capped = bool(flag) and max_rank is not None and max_rank < K
    else:
        capped = bool(flag)
    return rank, capped


def _truncate_core_right(core3: torch.Tensor, mode, error_threshold, max_rank, op=None, b_loc=None,
                         truncation_tol=None, residual_damp=2.0, stats: Optional[_SweepStats] = None,
                         spectra: Optional[list] = None):
    """SVD split (r_l n x r_r) -> (U_r, S V_r) (``amen.py:227-275``)."""
    rl, n, rr = core3.shape
    m = rl * n
    N = m * rr
    dec = _Decision()
    exact = D.empty(1)
    if mode is TruncationMode.RESIDUAL:
        tmp = D.empty(N)
        op.apply(P(core3), 1, P(tmp))
        D.diff_nrm2(N, P(tmp), P(b_loc), P(exact))
    U, S, Vt = D.svd(m, rr, P(core3), rr, 1, dec.buf, 0)
    if spectra is not None:
        spectra.append(S)
    K = min(m, rr)
    _choose_rank(mode, op, b_loc, U, S, Vt, m, rr, K, exact, error_threshold, max_rank,
                 truncation_tol, residual_damp, dec)
    rank, capped = _read_decision(dec, op, mode, max_rank, K)
    if stats is not None:
        stats.record_rank(capped)
    core = D.empty(rl, n, rank)
    D.copy_nd([m, rank], P(U), [K, 1], P(core), [rank, 1])
    carry = Vt[:rank].clone()
    D.scale_cols(rr, rank, P(carry), 1, rr, P(S))
    return core, carry


def _truncate_core_left(core3: torch.Tensor, mode, error_threshold, max_rank, op=None, b_loc=None,
                        truncation_tol=None, residual_damp=2.0, stats: Optional[_SweepStats] = None,
                        spectra: Optional[list] = None):
    """Mirror (``amen.py:278-322``): (V_r as (r, n, r_r), U_r S)."""
    rl, n, rr = core3.shape
    q = n * rr
    N = rl * q
    dec = _Decision()
    exact = D.empty(1)
    if mode is TruncationMode.RESIDUAL:
        tmp = D.empty(N)
        op.apply(P(core3), 1, P(tmp))
        D.diff_nrm2(N, P(tmp), P(b_loc), P(exact))
    U, S, Vt = D.svd(rl, q, P(core3), q, 1, dec.buf, 0)
    if spectra is not None:
        spectra.append(S)
    K = min(rl, q)
    _choose_rank(mode, op, b_loc, U, S, Vt, rl, q, K, exact, error_threshold, max_rank,
                 truncation_tol, residual_damp, dec)
    rank, capped = _read_decision(dec, op, mode, max_rank, K)
    if stats is not None:
        stats.record_rank(capped)
    core = Vt[:rank].reshape(rank, n, rr)
    carry = D.empty(rl, rank)
    D.copy_nd([rl, rank], P(U), [K, 1], P(carry), [rank, 1])
    D.scale_cols(rl, rank, P(carry), rank, 1, P(S))
    return core, carry


def _enrich_right(x: TensorTrain, k: int, core: torch.Tensor, carry: torch.Tensor,
                  block: torch.Tensor, max_rank=None, stats: Optional[_SweepStats] = None) -> None:
    """Append ``block`` on the right bond (``amen.py:325-360``)."""
    rl, n, rank = core.shape
    rho = block.shape[2]
    W = rank + rho
    enriched = D.empty(rl, n, W)
    D.copy_nd([rl * n, rank], P(core), [rank, 1], P(enriched), [W, 1])
    D.copy_nd([rl * n, rho], P(block), [rho, 1], P(enriched, rank), [W, 1])
    nxt = x.cores[k + 1]
    rr_old = nxt.shape[0]
    rest = nxt.numel() // rr_old
    newnext = D.empty((W,) + tuple(nxt.shape[1:]))
    D.gemm(1, rank, rest, rr_old, 1.0, P(carry), rr_old, 1, 0, P(nxt), rest, 1, 0, 0.0,
           P(newnext), rest, 1, 0)
    D.fill(newnext[rank:], 0.0)
    x.cores[k + 1] = newnext
    _shift_center_right(x, k, enriched)
    capped = max_rank is not None and x.bonds[k + 1] > max_rank
    if capped:
        x.cores[k] = _slice_last(x.cores[k], max_rank)
        x.cores[k + 1] = x.cores[k + 1][:max_rank]
        x.bonds[k + 1] = max_rank
    if stats is not None:
        stats.record_rank(capped)


def _enrich_left(x: TensorTrain, k: int, core: torch.Tensor, carry: torch.Tensor,
                 block: torch.Tensor, max_rank=None, stats: Optional[_SweepStats] = None) -> None:
    """Mirror (``amen.py:363-392``)."""
    rank, n, rr = core.shape
    rho = block.shape[0]
    W = rank + rho
    enriched = D.empty(W, n, rr)
    D.copy_nd([rank * n * rr], P(core), [1], P(enriched), [1])
    D.copy_nd([rho * n * rr], P(block), [1], P(enriched, rank * n * rr), [1])
    prev = x.cores[k - 1]
    rl_old = prev.shape[-1]
    rows = prev.numel() // rl_old
    newprev = D.zeros(tuple(prev.shape[:-1]) + (W,))
    D.gemm(1, rows, rank, rl_old, 1.0, P(prev), rl_old, 1, 0, P(carry), rank, 1, 0, 0.0,
           P(newprev), W, 1, 0)
    x.cores[k - 1] = newprev
    _shift_center_left(x, k, enriched)
    capped = max_rank is not None and x.bonds[k] > max_rank
    if capped:
        x.cores[k] = x.cores[k][:max_rank]
        x.cores[k - 1] = _slice_last(x.cores[k - 1], max_rank)
        x.bonds[k] = max_rank
    if stats is not None:
        stats.record_rank(capped)


def amen(A: TensorTrain, b: TensorTrain, x0: Optional[TensorTrain] = None, inplace: bool = False,
         kickrank: int = 4, error_threshold: float = 1e-7, max_rank: Optional[int] = None,
         max_sweeps: int = 1, residual_tol: Optional[float] = None, residual_damp: float = 2.0,
         stagnation_factor: Union[float, str, None] = "auto",
         local_solver: Union[str, LocalSolver] = "auto", direct_solve_size: int = 1000,
         local_rtol: Optional[float] = None, allow_dense_fallback: bool = False,
         symmetric: bool = False, return_info: bool = False, seed: Optional[int] = None
         ) -> Union[TensorTrain, Tuple[TensorTrain, SolverInfo]]:
    """Alternating Minimal Energy method for ``A x = b`` (``amen.py:395-725``)."""
    if kickrank <= 0:
        raise ValueError(f"kickrank must be a positive integer, got {kickrank}")
    if max_rank is not None and max_rank <= 0:
        raise ValueError(f"max_rank must be a positive integer, got {max_rank}")
    if residual_damp < 1.0:
        raise ValueError(f"residual_damp must be >= 1, got {residual_damp}")
    settings = _SolverSettings.resolve(
        max_sweeps=max_sweeps, residual_tol=residual_tol, stagnation_factor=stagnation_factor,
        local_solver=local_solver, symmetric=symmetric, direct_solve_size=direct_solve_size,
        return_info=return_info, local_rtol=local_rtol, allow_dense_fallback=allow_dense_fallback,
        truncation_steps=max(2 * (A.n_cores - 1), 1))
    if x0 is None:
        x0 = ones(b.physical_dims(), ranks=b.bonds)
    _validate_solver_inputs(A, b, x0)
    x = x0 if inplace else x0.copy()
    if x.orth_center != 0:
        x.orthonormalize(canonical="right", inplace=True)

    d = len(x)
    mode = settings.truncation_mode
    residual_truncation = mode is TruncationMode.RESIDUAL
    rA, rb, lA, lb = [None] * d, [None] * d, [None] * d, [None] * d
    z = init_residual_train(b, kickrank, seed=seed)
    fit_residual_train(A, b, x, z, sweeps=1)
    lzb, rzb, lzA, rzA, lxz, rxz = [None] * d, [None] * d, [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, -1, -1):
        prepare_right_environment(i, rA, x, x, A)
        prepare_right_environment(i, rb, x, b)
        prepare_right_environment(i, rzb, z, b)
        prepare_right_environment(i, rzA, z, x, A)
        prepare_right_environment(i, rxz, x, z)
    monitor = _SweepMonitor(settings, "AMEn", A, b)
    stats = _SweepStats(max_rank=max_rank)
    tt = settings.truncation_tol

    for sweep in range(settings.max_sweeps):
        for k in range(d - 1):
            prepare_left_environment(k, lA, x, x, A)
            prepare_left_environment(k, lb, x, b)
            prepare_left_environment(k, lzb, z, b)
            prepare_left_environment(k, lzA, z, x, A)
            prepare_left_environment(k, lxz, x, z)
            u, op, b_loc = _solve_local_one_site(k, x, A, b, lA, rA, lb, rb, settings,
                                                 preserve_operator=residual_truncation, stats=stats)
            core3 = u.reshape(x.bonds[k], x.cores[k].shape[1], x.bonds[k + 1])
            z_new = residual_core(k, lzb, rzb, lzA, rzA, A, b, x, x_core=core3)
            core, carry = _truncate_core_right(core3, mode, error_threshold, max_rank, op, b_loc, tt,
                                               residual_damp, stats=stats)
            xl = lxz[k].shape[1]
            zl, nz, zr = z_new.shape
            block = D.empty(xl, nz, zr)
            D.gemm(1, xl, nz * zr, zl, 1.0, P(lxz[k]), 1, xl, 0, P(z_new), nz * zr, 1, 0, 0.0,
                   P(block), nz * zr, 1, 0)
            _enrich_right(x, k, core, carry, block, max_rank, stats=stats)
            _shift_center_right(z, k, z_new)

        prepare_left_environment(d - 1, lA, x, x, A)
        prepare_left_environment(d - 1, lb, x, b)
        prepare_left_environment(d - 1, lzb, z, b)
        prepare_left_environment(d - 1, lzA, z, x, A)
        prepare_left_environment(d - 1, lxz, x, z)
        if _log.isEnabledFor(logging.DEBUG):
            _log.debug("AMEn sweep %d: forward half-sweep done, max bond %d", sweep + 1, max(x.bonds))

        for k in range(d - 1, -1, -1):
            prepare_right_environment(k, rA, x, x, A)
            prepare_right_environment(k, rb, x, b)
            prepare_right_environment(k, rzb, z, b)
            prepare_right_environment(k, rzA, z, x, A)
            prepare_right_environment(k, rxz, x, z)
            u, op, b_loc = _solve_local_one_site(k, x, A, b, lA, rA, lb, rb, settings,
                                                 preserve_operator=residual_truncation, stats=stats)
            core3 = u.reshape(x.bonds[k], x.cores[k].shape[1], x.bonds[k + 1])
            z_new = residual_core(k, lzb, rzb, lzA, rzA, A, b, x, x_core=core3)
            if k == 0:
                x.cores[0] = core3.reshape(x.bonds[0], core3.shape[1], 1, x.bonds[1])
                x.orth_center = 0
                z.cores[0] = z_new.reshape(z.bonds[0], z_new.shape[1], 1, z.bonds[1])
                z.orth_center = 0
            else:
                core, carry = _truncate_core_left(core3, mode, error_threshold, max_rank, op, b_loc,
                                                  tt, residual_damp, stats=stats)
                block = D.absorb_right(z_new, P(rxz[k]), rxz[k].shape[0], rxz[k].shape[1])
                _enrich_left(x, k, core, carry, block, max_rank, stats=stats)
                _shift_center_left(z, k, z_new)

        if _log.isEnabledFor(logging.DEBUG):
            _log.debug("AMEn sweep %d: backward half-sweep done, max bond %d; residual check",
                       sweep + 1, max(x.bonds))
        if monitor.record(sweep, x, max_bond=max(x.bonds), stats=stats):
            break

    redundant = [k for k in range(1, d) if x.bonds[k] > prod(x.row_modes[:k])]
    if redundant:
        x.orthonormalize(canonical="left", center_position=redundant[-1], inplace=True)
        x.orthonormalize(canonical="right", center_position=0, inplace=True)
    x.max_rank = max(x.bonds)
    monitor.finish()
    if return_info:
        return x, monitor.info
    return x
