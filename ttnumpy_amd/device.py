"""Device operations of the solver hot path, issued to libttx on torch buffers.

PyTorch provides device memory and the stream; every arithmetic operation is
a libttx kernel.  Cores are contiguous FP64 device tensors
``(left_bond, row_mode, col_mode, right_bond)``; environments are contiguous
``(target, operator, frame)`` or ``(target, frame)`` tensors, the index order
of ``ttnumpy/solvers/common.py:149-217``.

Contractions are issued as strided batched GEMMs whose strides read the
operands in place, so no permuted copy of a core or environment is made.
"""

from __future__ import annotations

from typing import List, Optional, Tuple

import torch

from . import _lib
from ._lib import check

F64 = torch.float64
_DEV = None


def device() -> torch.device:
    global _DEV
    if _DEV is None:
        if not torch.cuda.is_available():
            raise RuntimeError("ttnumpy_amd needs a HIP device (MI355X); none is visible")
        _lib.load()
        _DEV = torch.device("cuda", torch.cuda.current_device())
    return _DEV


def lib():
    return _lib.load()


_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream() -> int:
    """Raw handle of torch's current stream on the current device (all ttx work goes there)."""
    if _RAW_STREAM is not None:
        return _RAW_STREAM(device().index)
    return torch.cuda.current_stream().cuda_stream


def empty(*shape) -> torch.Tensor:
    return torch.empty(*shape, dtype=F64, device=device())


def zeros(*shape) -> torch.Tensor:
    t = torch.empty(*shape, dtype=F64, device=device())
    fill(t, 0.0)
    return t


def to_device(a) -> torch.Tensor:
    return torch.as_tensor(a, dtype=F64).to(device()).contiguous()


def P(t: torch.Tensor, off: int = 0) -> int:
    """Device address of element ``off`` of a contiguous FP64 tensor."""
    return t.data_ptr() + 8 * off


# ---------------------------------------------------------------- flop tally
# Algorithmic flops of the launches made through this module, counted on the
# host from each call's shapes with the closed forms of the routines (GEMM
# 2mnk; Householder QR 2mnk - (m+n)k^2 + 2k^3/3 with Q formed as
# 4mk^2 - 4k^3/3; SVD as the QR of the long side plus 20 Q^3 + 2 P Q^2; LU
# 2n^3/3 + 2n^2; Cholesky n^3/3 + 2n^2).  ``flops()`` reads the running total.
_FLOPS = [0.0]


def add_flops(f: float) -> None:
    _FLOPS[0] += f


def flops() -> float:
    return _FLOPS[0]


def _qr_flops(m: int, n: int, form_q: bool) -> float:
    k = min(m, n)
    f = 2.0 * m * n * k - (m + n) * k * k + 2.0 * k ** 3 / 3.0
    return f + (4.0 * m * k * k - 4.0 * k ** 3 / 3.0 if form_q else 0.0)


# ---------------------------------------------------------------- primitives
# While a recorder dict is installed here, the FP64 GEMMs issued through this
# module are counted by shape: (batch, M, N, K, A m-fastest, B n-fastest) -> calls
# (ttnumpy_amd.autotune times those shapes).
_GEMM_RECORD: List[Optional[dict]] = [None]


def record_gemm(batch, M, N, K, a_m_fast: bool, b_n_fast: bool, count: int = 1) -> None:
    rec = _GEMM_RECORD[0]
    if rec is not None:
        key = (int(batch), int(M), int(N), int(K), bool(a_m_fast), bool(b_n_fast))
        rec[key] = rec.get(key, 0) + int(count)


def gemm(batch, M, N, K, alpha, A, a_ms, a_ks, a_bs, B, b_ks, b_ns, b_bs, beta, C, c_ms, c_ns, c_bs):
    """C = alpha A B + beta C on MFMA; A/B/C are addresses (ints)."""
    _FLOPS[0] += 2.0 * batch * M * N * K
    if _GEMM_RECORD[0] is not None:
        record_gemm(batch, M, N, K, a_ms <= a_ks, b_ns <= b_ks)
    rc = lib().ttx_gemm_f64(stream(), batch, M, N, K, alpha, A, a_ms, a_ks, a_bs, B, b_ks, b_ns,
                            b_bs, beta, C, c_ms, c_ns, c_bs)
    check(rc, "ttx_gemm_f64")


# Precision of the absorb contractions (``absorb_left``/``absorb_right``):
# "fp64", or "fp32" for the mixed rounding form (operands rounded to FP32,
# FP32 accumulation on MFMA, FP64 storage).  Set through ``contractions``.
_CONTRACT = ["fp64"]


class contractions:
    """Context manager selecting the absorb-contraction precision."""

    def __init__(self, precision: str):
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"contraction precision must be 'fp64' or 'fp32', got {precision!r}")
        self.precision = precision

    def __enter__(self):
        self.saved = _CONTRACT[0]
        _CONTRACT[0] = self.precision
        return self

    def __exit__(self, *exc):
        _CONTRACT[0] = self.saved
        return False


def gemm_contract(batch, M, N, K, alpha, A, a_ms, a_ks, a_bs, B, b_ks, b_ns, b_bs, beta, C, c_ms,
                  c_ns, c_bs, a_count: int = 0, b_count: int = 0):
    """``gemm`` in the precision selected by ``contractions``.  In FP32 the
    operands are scaled by their norms over their ``a_count``/``b_count``
    contiguous elements (0: unscaled) so the rounding stays in FP32's range."""
    if _CONTRACT[0] == "fp64":
        return gemm(batch, M, N, K, alpha, A, a_ms, a_ks, a_bs, B, b_ks, b_ns, b_bs, beta, C, c_ms,
                    c_ns, c_bs)
    _FLOPS[0] += 2.0 * batch * M * N * K
    scales = empty(2)
    sa = sb = None
    if a_count:
        check(lib().ttx_nrm2_f64(stream(), a_count, A, P(scales)), "ttx_nrm2_f64")
        sa = P(scales)
    if b_count:
        check(lib().ttx_nrm2_f64(stream(), b_count, B, P(scales, 1)), "ttx_nrm2_f64")
        sb = P(scales, 1)
    rc = lib().ttx_gemm_f32acc(stream(), batch, M, N, K, alpha, A, a_ms, a_ks, a_bs, B, b_ks, b_ns,
                               b_bs, beta, C, c_ms, c_ns, c_bs, sa, sb)
    check(rc, "ttx_gemm_f32acc")
    _KEEP_SCALES.append(scales)
    if len(_KEEP_SCALES) > 64:
        del _KEEP_SCALES[:32]


_KEEP_SCALES = []  # scale buffers stay alive until their kernels have long been enqueued


def copy_nd(dims, src, src_strides, dst, dst_strides, alpha=1.0):
    nd = len(dims)
    arr = (ctypes_i64 * nd)
    rc = lib().ttx_copy_f64(stream(), nd, arr(*dims), src, arr(*src_strides), dst,
                            arr(*dst_strides), alpha)
    check(rc, "ttx_copy_f64")


def fill(t: torch.Tensor, value: float) -> None:
    check(lib().ttx_fill_f64(stream(), P(t), t.numel(), value), "ttx_fill_f64")


def nrm2(n: int, x: int, out: int) -> None:
    check(lib().ttx_nrm2_f64(stream(), n, x, out), "ttx_nrm2_f64")


def diff_nrm2(n: int, a: int, b: int, out: int) -> None:
    check(lib().ttx_diff_nrm2_f64(stream(), n, a, b, out), "ttx_diff_nrm2_f64")


def norm(t: torch.Tensor) -> torch.Tensor:
    """Device scalar ||t||_2."""
    out = empty(1)
    nrm2(t.numel(), P(t), P(out))
    return out


import ctypes as _ct  # noqa: E402

ctypes_i64 = _ct.c_int64


# ---------------------------------------------------------------- environments
def env_left(E: torch.Tensor, F: torch.Tensor, T: torch.Tensor, A: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Next left environment (``common.py:170-188``); F frame core, T target core (4-D)."""
    rf, nr, _, rf2 = F.shape
    rt, nc, _, rt2 = T.shape
    if A is None:
        s1 = empty(rt * nr * rf2)
        gemm(1, rt, nr * rf2, rf, 1.0, P(E), rf, 1, 0, P(F), nr * rf2, 1, 0, 0.0, P(s1), nr * rf2, 1, 0)
        out = empty(rt2, rf2)
        gemm(1, rt2, rf2, rt * nr, 1.0, P(T), 1, rt2, 0, P(s1), rf2, 1, 0, 0.0, P(out), rf2, 1, 0)
        return out
    Ra, _, _, Ra2 = A.shape
    s1 = empty(rt * Ra * nr * rf2)       # [t, a, i, f']
    gemm(1, rt * Ra, nr * rf2, rf, 1.0, P(E), rf, 1, 0, P(F), nr * rf2, 1, 0, 0.0, P(s1), nr * rf2, 1, 0)
    s2 = empty(rt * nc * Ra2 * rf2)      # [t, j, a', f']
    gemm(rt, rf2, nc * Ra2, Ra * nr, 1.0, P(s1), 1, rf2, Ra * nr * rf2, P(A), nc * Ra2, 1, 0, 0.0,
         P(s2), 1, rf2, nc * Ra2 * rf2)
    out = empty(rt2, Ra2, rf2)
    gemm(1, rt2, Ra2 * rf2, rt * nc, 1.0, P(T), 1, rt2, 0, P(s2), Ra2 * rf2, 1, 0, 0.0, P(out),
         Ra2 * rf2, 1, 0)
    return out


def env_right(Phi: torch.Tensor, F: torch.Tensor, T: torch.Tensor, A: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Next right environment (``common.py:203-217``); F, T are the cores right of it."""
    rf, nr, _, rf2 = F.shape
    rt, nc, _, rt2 = T.shape
    if A is None:
        s1 = empty(rt * nc * rf2)        # [t, i, f']
        gemm(1, rt * nc, rf2, rt2, 1.0, P(T), rt2, 1, 0, P(Phi), rf2, 1, 0, 0.0, P(s1), rf2, 1, 0)
        out = empty(rt, rf)
        gemm(1, rt, rf, nc * rf2, 1.0, P(s1), nc * rf2, 1, 0, P(F), 1, nr * rf2, 0, 0.0, P(out), rf, 1, 0)
        return out
    Ra, _, _, Ra2 = A.shape
    s1 = empty(rt * nc * Ra2 * rf2)      # [t, j, a', f']
    gemm(1, rt * nc, Ra2 * rf2, rt2, 1.0, P(T), rt2, 1, 0, P(Phi), Ra2 * rf2, 1, 0, 0.0, P(s1),
         Ra2 * rf2, 1, 0)
    s2 = empty(rt * Ra * nr * rf2)       # [t, a, i, f']
    gemm(rt, rf2, Ra * nr, nc * Ra2, 1.0, P(s1), 1, rf2, nc * Ra2 * rf2, P(A), 1, nc * Ra2, 0, 0.0,
         P(s2), 1, rf2, Ra * nr * rf2)
    out = empty(rt, Ra, rf)
    gemm(1, rt * Ra, rf, nr * rf2, 1.0, P(s2), nr * rf2, 1, 0, P(F), 1, nr * rf2, 0, 0.0, P(out), rf, 1, 0)
    return out


def ones_env(with_op: bool) -> torch.Tensor:
    t = empty(1, 1, 1) if with_op else empty(1, 1)
    fill(t, 1.0)
    return t


# ---------------------------------------------------------------- local system
class LocalOp:
    """Matrix-free one-site operator Psi . A_k . Phi (``one_site.py:78-115``).

    Rows are indexed by the frame bonds (fl, i, fr), columns by the target
    bonds (rl, j, rr).  ``apply`` takes B column vectors stored one after
    the other (column-major N x B).
    """

    def __init__(self, psi: torch.Tensor, A: torch.Tensor, phi: torch.Tensor):
        self.psi, self.A, self.phi = psi, A, phi
        self.rl, self.Ra, self.fl = psi.shape
        self.rr, self.Rb, self.fr = phi.shape
        _, self.nr, self.nc, _ = A.shape
        self.rows = self.fl * self.nr * self.fr
        self.cols = self.rl * self.nc * self.rr
        self._scratch = {}

    def _bufs(self, B):
        if B not in self._scratch:
            self._scratch[B] = (empty(B * self.rl * self.nc * self.Rb * self.fr),
                                empty(B * self.rl * self.Ra * self.nr * self.fr))
        return self._scratch[B]

    def flops_per_apply(self, B: int = 1) -> float:
        rl, Ra, fl, rr, Rb, fr, nr, nc = self.rl, self.Ra, self.fl, self.rr, self.Rb, self.fr, self.nr, self.nc
        return 2.0 * B * (rl * nc * Rb * fr * rr + rl * Ra * nr * fr * nc * Rb + fl * nr * fr * rl * Ra)

    def _record(self, B: int, count: int) -> None:
        rl, Ra, fl, rr, Rb, fr, nr, nc = self.rl, self.Ra, self.fl, self.rr, self.Rb, self.fr, self.nr, self.nc
        record_gemm(1, B * rl * nc, Rb * fr, rr, rr <= 1, True, count)
        record_gemm(B * rl, Ra * nr, fr, nc * Rb, nc * Rb <= 1, True, count)
        record_gemm(B, fl, nr * fr, rl * Ra, True, True, count)

    def apply(self, y: int, B: int, out: int, alpha: float = 1.0, beta: float = 0.0) -> None:
        """s1 = y . phi, s2 = A_k . s1, out = psi^T . s2 (``ttx_local_apply_f64``)."""
        s1, s2 = self._bufs(B)
        _FLOPS[0] += self.flops_per_apply(B)
        if _GEMM_RECORD[0] is not None:
            self._record(B, 1)
        rc = lib().ttx_local_apply_f64(stream(), P(self.psi), self.rl, self.Ra, self.fl, P(self.A),
                                       self.nr, self.nc, self.Rb, P(self.phi), self.rr, self.fr, y,
                                       B, alpha, beta, out, P(s1), P(s2))
        check(rc, "ttx_local_apply_f64")

    def gmres_cycle(self, m: int, b: int, work: int, st: int) -> None:
        """One GMRES cycle with this operator (``ttx_gmres_local_cycle_f64``); the
        caller tallies the flops of the steps that ran."""
        s1, s2 = self._bufs(1)
        if _GEMM_RECORD[0] is not None:
            self._record(1, m + 1)
        rc = lib().ttx_gmres_local_cycle_f64(stream(), P(self.psi), self.rl, self.Ra, self.fl,
                                             P(self.A), self.nr, self.nc, self.Rb, P(self.phi),
                                             self.rr, self.fr, P(s1), P(s2), m, b, work, st)
        check(rc, "ttx_gmres_local_cycle_f64")

    def minres_steps(self, y_in: int, vec: int, st: int, count: int) -> None:
        """``count`` MINRES iterations with this operator (``ttx_minres_local_steps_f64``);
        the caller tallies the flops of the iterations that ran."""
        s1, s2 = self._bufs(1)
        if _GEMM_RECORD[0] is not None:
            self._record(1, count)
        rc = lib().ttx_minres_local_steps_f64(stream(), P(self.psi), self.rl, self.Ra, self.fl,
                                              P(self.A), self.nr, self.nc, self.Rb, P(self.phi),
                                              self.rr, self.fr, P(s1), P(s2), y_in, vec, st, count)
        check(rc, "ttx_minres_local_steps_f64")

    def explicit(self, out: torch.Tensor, ld: int) -> None:
        """Write the explicit matrix row-major (row stride ``ld``) into ``out`` (``one_site.py:49-75``)."""
        rc = lib().ttx_local_matrix_f64(stream(), P(self.psi), self.rl, self.Ra, self.fl, P(self.A),
                                        self.nr, self.nc, self.Rb, P(self.phi), self.rr, self.fr,
                                        P(out), ld, 1)
        check(rc, "ttx_local_matrix_f64")


class ExplicitOp:
    """The explicit local matrix (row-major N x N) as an operator."""

    def __init__(self, M: torch.Tensor, n: int):
        self.M, self.rows, self.cols = M, n, n

    def apply(self, y: int, B: int, out: int, alpha: float = 1.0, beta: float = 0.0) -> None:
        n = self.rows
        gemm(1, n, B, n, alpha, P(self.M), n, 1, 0, y, 1, n, 0, beta, out, 1, n, 0)


def local_rhs(psi_b: torch.Tensor, b_core: torch.Tensor, phi_b: torch.Tensor, out: int) -> None:
    """``b_loc[f_l, i, f_r] = sum Psi[t, f_l] b[t, i, t'] Phi[t', f_r]`` (``one_site.py:27-46``)."""
    rt, fl = psi_b.shape
    _, n, _, rt2 = b_core.shape
    _, fr = phi_b.shape
    s1 = empty(fl * n * rt2)
    gemm(1, fl, n * rt2, rt, 1.0, P(psi_b), 1, fl, 0, P(b_core), n * rt2, 1, 0, 0.0, P(s1), n * rt2, 1, 0)
    gemm(1, fl * n, fr, rt2, 1.0, P(s1), rt2, 1, 0, P(phi_b), fr, 1, 0, 0.0, out, fr, 1, 0)


def lu_solve(Mb: torch.Tensor, n: int, u: torch.Tensor, info: torch.Tensor, ld: Optional[int] = None) -> None:
    """Solve the row-major [M | b] system in ``Mb`` (destroyed) into ``u`` (``common.py:767``)."""
    _FLOPS[0] += 2.0 * n ** 3 / 3.0 + 2.0 * n * n
    work = torch.empty(int(lib().ttx_lu_work_size(n)) + 1, dtype=F64, device=device())
    check(lib().ttx_lu_solve_f64(stream(), n, P(Mb), n + 1 if ld is None else ld, P(u), P(work),
                                 info.data_ptr()), "ttx_lu_solve_f64")


def chol_solve(Mb: torch.Tensor, n: int, u: torch.Tensor, info: torch.Tensor,
               ld: Optional[int] = None) -> None:
    """Cholesky solve of the SPD row-major [M | b] in ``Mb`` (lower triangle read,
    destroyed); ``info`` gets TTX_ENOTSPD when M is not positive definite."""
    _FLOPS[0] += n ** 3 / 3.0 + 2.0 * n * n
    check(lib().ttx_chol_solve_f64(stream(), n, P(Mb), n + 1 if ld is None else ld, P(u),
                                   info.data_ptr()), "ttx_chol_solve_f64")


# ---------------------------------------------------------------- factorizations
def qr(m: int, n: int, A: int, a_rs: int, a_cs: int, Q: Optional[int], q_rs: int, q_cs: int,
       R: int, r_rs: int, r_cs: int) -> None:
    _FLOPS[0] += _qr_flops(m, n, Q is not None)
    work = empty(int(lib().ttx_qr_work_size(m, n)))
    rc = lib().ttx_qr_f64(stream(), m, n, A, a_rs, a_cs, Q if Q is not None else None, q_rs, q_cs,
                          R, r_rs, r_cs, P(work))
    check(rc, "ttx_qr_f64")


def svd(m: int, n: int, A: int, a_rs: int, a_cs: int, info: torch.Tensor, info_off: int = 0
        ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Thin SVD; returns contiguous U (m x k), S (k), Vt (k x n)."""
    k = min(m, n)
    lp, sq = max(m, n), k
    _FLOPS[0] += (_qr_flops(lp, sq, True) if lp > sq else 0.0) + 20.0 * sq ** 3 + 2.0 * lp * sq * sq
    U = empty(m, k)
    S = empty(k)
    Vt = empty(k, n)
    work = empty(int(lib().ttx_svd_work_size(m, n)))
    rc = lib().ttx_svd_f64(stream(), m, n, A, a_rs, a_cs, P(U), k, 1, P(S), P(Vt), n, 1, P(work),
                           info.data_ptr() + 4 * info_off)
    check(rc, "ttx_svd_f64")
    return U, S, Vt


def scale_cols(rows, cols, X: int, rs, cs, s: int) -> None:
    check(lib().ttx_scale_cols_f64(stream(), rows, cols, X, rs, cs, s), "ttx_scale_cols_f64")


def truncated_bond_dev(S: torch.Tensor, nrm: torch.Tensor, scale: float, max_rank: Optional[int],
                       out: torch.Tensor, out_off: int = 0) -> None:
    rc = lib().ttx_truncated_bond_dev_f64(stream(), P(S), S.numel(), P(nrm), scale,
                                          0 if max_rank is None else int(max_rank),
                                          out.data_ptr() + 4 * out_off)
    check(rc, "ttx_truncated_bond_dev_f64")


def ttm_core(A: torch.Tensor, x: torch.Tensor, out: int, o_l: int, o_i: int, o_r: int) -> None:
    Ra, nr, nc, Rb = A.shape
    rx, _, _, rx2 = x.shape
    rc = lib().ttx_ttm_core_f64(stream(), P(A), Ra, nr, nc, Rb, P(x), rx, rx2, out, o_l, o_i, o_r)
    check(rc, "ttx_ttm_core_f64")


def absorb_left(Mx: int, m: int, r: int, core: torch.Tensor, out: Optional[torch.Tensor] = None,
                m_rs: Optional[int] = None, m_cs: int = 1) -> torch.Tensor:
    """``(m x r) @ core.reshape(r, -1)`` (``utils.py:72-76``); Mx row-major unless strides given."""
    rest = core.numel() // core.shape[0]
    if out is None:
        out = empty((m,) + tuple(core.shape[1:]))
    gemm_contract(1, m, rest, r, 1.0, Mx, r if m_rs is None else m_rs, m_cs, 0, P(core), rest, 1,
                  0, 0.0, P(out), rest, 1, 0, a_count=m * r, b_count=core.numel())
    return out


def absorb_right(core: torch.Tensor, Mx: int, r: int, m: int, m_rs: Optional[int] = None,
                 m_cs: int = 1) -> torch.Tensor:
    """``core.reshape(-1, r) @ (r x m)`` (``utils.py:79-83``)."""
    rows = core.numel() // r
    out = empty(tuple(core.shape[:-1]) + (m,))
    gemm_contract(1, rows, m, r, 1.0, P(core), r, 1, 0, Mx, m if m_rs is None else m_rs, m_cs, 0,
                  0.0, P(out), m, 1, 0, a_count=core.numel(), b_count=r * m)
    return out


def sync_ints(t: torch.Tensor):
    """Bring a small int32 device array to the host (the one sync of a decision)."""
    return t.cpu().tolist()
