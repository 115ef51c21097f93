"""CPU oracle of the tensor-train solver hot path (test infrastructure only).

This module is the checker for the device implementation in ``ttnumpy_amd``:
only the tests, the smoke check and the benchmark's CPU-baseline leg import
it.  The product path never calls into it.

It restates, in FP64 NumPy/SciPy, the algorithm of ttnumpy 0.3.0 for

* ``TensorTrain.truncate`` -- QR sweep then truncated-SVD sweep
  (``ttnumpy/tensor_train.py:705-940``, ``ttnumpy/utils.py:22-136,336-460``);
* the one-site AMEn and ALS solvers (``ttnumpy/solvers/amen.py:80-725``,
  ``ttnumpy/solvers/als.py:56-212``, ``ttnumpy/solvers/one_site.py:27-209``,
  ``ttnumpy/solvers/common.py:118-929``) and the two-site MALS
  (``ttnumpy/solvers/mals.py:61-292``, ``ttnumpy/solvers/two_site.py:28-241``);
* the QTT conversions ``tt_to_qtt`` / ``qtt_to_tt``
  (``ttnumpy/tensor_train.py:58-91,942-1135``);
* the per-sweep residual ``||A x - b|| / ||b||`` in TT arithmetic
  (``common.py:118-146`` via ``tensor_train.py:254-540``).

Every core is a 4-D array ``(left_bond, row_mode, col_mode, right_bond)``; a
plain train has ``col_mode == 1``.  Trains are :class:`Train` records.

Pinning (what checks each function, beyond retyping its formula):

* ``truncated_bond``: the known-answer values of the reference's tests
  (``tests/test_utils.py:57-130``) -- see ``tests/test_oracle.py``.
* ``truncate``: golden fixtures written by the reference itself
  (``tests/golden/c2_round_d40.npz``, ``round_small_*.npz``), the two-core
  closed form ``error = ||sigma[2:]||/||sigma||`` and orthonormality checks;
  ``method="randomized"``: the reference's fixtures ``round_rand_*.npz``
  (sketches below and above the train's ranks, both sweep directions) and
  the exactness of the sketch when no unfolding has rank above its bond.
* environments / local operator / local rhs: brute force against dense
  interface matrices on tiny trains (``x^T A x`` projections).
* ``tt_to_qtt`` / ``qtt_to_tt`` / ``merge_cores``: the reference's fixtures
  ``qtt_*.npz`` (truncating plain trains, a TTM with a rank cap, mode 3,
  regrouping), the properties its own tests assert (round trips, error
  bound, canonical form, bond floor, exact ranks) and, for the merge, dense
  Kronecker products.
* ``tt_multiply`` / ``randomized_hadamard``: the dense element-wise product,
  the reference's fixture ``multiply_rand.npz``.
* ``mals``: the reference's trajectories ``mals_*.npz`` and a brute-force
  projection of the two-site system.
* ``amen`` / ``als`` / ``relative_residual``: the golden trajectories written
  by the reference (bonds, residual history, stop reason, solution), and the
  dense solution of small systems.
"""

from __future__ import annotations

import numbers
from dataclasses import dataclass, field
from math import prod
from typing import List, Optional, Tuple

import numpy as np
import scipy.linalg as sla
import scipy.sparse.linalg as spla


# --------------------------------------------------------------------------
# container
# --------------------------------------------------------------------------


@dataclass
class Train:
    """List of 4-D cores plus the canonical-centre bookkeeping.

    Mirrors the fields of the reference container that the hot path reads
    (``tensor_train.py:150-244``): ``bonds`` (d+1 ints), ``orth_center``.
    """

    cores: List[np.ndarray]
    orth_center: Optional[int] = None

    def __post_init__(self):
        cores = []
        for c in self.cores:
            c = np.asarray(c, dtype=np.float64)
            if c.ndim == 3:
                c = c.reshape(c.shape[0], c.shape[1], 1, c.shape[2])
            cores.append(c)
        self.cores = cores

    @property
    def d(self) -> int:
        return len(self.cores)

    @property
    def bonds(self) -> List[int]:
        return [c.shape[0] for c in self.cores] + [self.cores[-1].shape[3]]

    @property
    def row_modes(self) -> List[int]:
        return [c.shape[1] for c in self.cores]

    @property
    def col_modes(self) -> List[int]:
        return [c.shape[2] for c in self.cores]

    def copy(self) -> "Train":
        return Train([c.copy() for c in self.cores], self.orth_center)

    def full(self) -> np.ndarray:
        """Dense tensor (row modes then col modes interleaved per core), small trains only."""
        t = self.cores[0].reshape(-1, self.cores[0].shape[3])
        for c in self.cores[1:]:
            t = (t @ c.reshape(c.shape[0], -1)).reshape(-1, c.shape[3])
        return t.reshape(-1)

    def dense_matrix(self) -> np.ndarray:
        """Dense matrix of a TTM (rows = row multi-index, cols = col multi-index)."""
        rows, cols = self.row_modes, self.col_modes
        d = self.d
        t = self.full().reshape([m for pair in zip(rows, cols) for m in pair])
        perm = list(range(0, 2 * d, 2)) + list(range(1, 2 * d, 2))
        return t.transpose(perm).reshape(prod(rows), prod(cols))


# --------------------------------------------------------------------------
# rounding: QR sweeps, SVD sweeps, truncate  (utils.py, tensor_train.py)
# --------------------------------------------------------------------------


def truncated_bond(s: np.ndarray, delta: float, max_rank: Optional[int] = None) -> int:
    """Smallest rank whose discarded tail has Frobenius norm <= delta.

    ``utils.py:22-69``: values <= ``len(s) * eps * s[0]`` count as zero, the
    tail is the reversed cumulative sum of squares, floor 1, cap max_rank.
    """
    if max_rank is not None and max_rank <= 0:
        raise ValueError("max_rank must be a positive integer.")
    s = np.asarray(s)
    eps = np.finfo(s.dtype).eps if np.issubdtype(s.dtype, np.inexact) else np.finfo(np.float64).eps
    v = s.astype(float).copy()
    if v.size:
        v[v <= eps * v.size * v[0]] = 0.0
    tail = np.cumsum((v ** 2)[::-1])[::-1]
    r = max(1, int(np.count_nonzero(tail > delta * delta)))
    if max_rank is not None:
        r = min(r, max_rank)
    return r


def _absorb_left(m: np.ndarray, core: np.ndarray) -> np.ndarray:
    """``m @ core`` over the core's first axis (``utils.py:72-76``)."""
    return (m @ core.reshape(core.shape[0], -1)).reshape((m.shape[0],) + core.shape[1:])


def _absorb_right(core: np.ndarray, m: np.ndarray) -> np.ndarray:
    """``core @ m`` over the core's last axis (``utils.py:79-83``)."""
    return (core.reshape(-1, core.shape[-1]) @ m).reshape(core.shape[:-1] + (m.shape[1],))


def qr_sweep_left(cores: List[np.ndarray], start: int, end: int) -> None:
    """Left-orthonormalise cores ``start..end-1`` in place (``utils.py:86-112``)."""
    for i in range(start, end):
        s = cores[i].shape
        q, r = sla.qr(cores[i].reshape(s[0] * s[1] * s[2], s[3]), mode="economic")
        cores[i] = q.reshape(s[0], s[1], s[2], -1)
        cores[i + 1] = _absorb_left(r, cores[i + 1])


def qr_sweep_right(cores: List[np.ndarray], start: int, end: int) -> None:
    """Right-orthonormalise cores ``start..end+1`` in place (``utils.py:115-136``)."""
    for i in range(start, end, -1):
        s = cores[i].shape
        q, r = sla.qr(cores[i].reshape(s[0], s[1] * s[2] * s[3]).T, mode="economic")
        cores[i] = q.T.reshape(-1, s[1], s[2], s[3])
        cores[i - 1] = _absorb_right(cores[i - 1], r.T)


def _thin_svd(m: np.ndarray, core: int):
    """``gesdd`` with a ``gesvd`` retry (``utils.py:336-361``)."""
    try:
        return sla.svd(m, full_matrices=False, check_finite=False)
    except np.linalg.LinAlgError as err:
        if not np.isfinite(m).all():
            raise ValueError(f"Core {core} has NaN or infinite entries") from err
        return sla.svd(m, full_matrices=False, check_finite=False, lapack_driver="gesvd")


def svd_sweep_left(cores, error, max_rank=None, spectra=None):
    """Left-to-right truncated SVD sweep (``utils.py:364-411``)."""
    c0 = cores[0]
    nrm = np.linalg.norm(c0.reshape(-1))
    delta = error / np.sqrt(len(cores) - 1) * nrm
    bond = 1
    for i in range(len(cores) - 1):
        s = cores[i].shape
        u, sv, vt = _thin_svd(cores[i].reshape(bond * s[1] * s[2], s[3]), i)
        if spectra is not None:
            spectra.append(sv.copy())
        nb = truncated_bond(sv, delta, max_rank)
        cores[i] = u[:, :nb].reshape(bond, s[1], s[2], nb)
        cores[i + 1] = _absorb_left(sv[:nb, None] * vt[:nb], cores[i + 1])
        bond = nb
    return cores


def svd_sweep_right(cores, error, max_rank=None, spectra=None):
    """Right-to-left truncated SVD sweep (``utils.py:414-460``)."""
    cl = cores[-1]
    nrm = np.linalg.norm(cl.reshape(-1))
    delta = error / np.sqrt(len(cores) - 1) * nrm
    bond = 1
    for i in range(len(cores) - 1, 0, -1):
        s = cores[i].shape
        u, sv, vt = _thin_svd(cores[i].reshape(s[0], s[1] * s[2] * bond), i)
        if spectra is not None:
            spectra.append(sv.copy())
        nb = truncated_bond(sv, delta, max_rank)
        cores[i] = vt[:nb].reshape(nb, s[1], s[2], bond)
        cores[i - 1] = _absorb_right(cores[i - 1], u[:, :nb] * sv[:nb])
        bond = nb
    return cores


def orthonormalize(t: Train, canonical: str = "left", center: Optional[int] = None) -> Train:
    """In-place QR sweep with the centre bookkeeping of ``tensor_train.py:705-787``."""
    d = t.d
    if center is not None and not 0 <= center < d:
        raise ValueError(f"center_position must be in [0, {d - 1}], but got {center}")
    had = t.orth_center is not None
    if canonical == "left":
        end = center if center is not None else d - 1
        start = 0
        if t.orth_center is not None:
            start = t.orth_center
            t.orth_center = end if start <= end else t.orth_center
        if start >= end:
            return t
        qr_sweep_left(t.cores, start, end)
    elif canonical == "right":
        end = center if center is not None else 0
        start = d - 1
        if t.orth_center is not None:
            start = t.orth_center
            t.orth_center = end if start >= end else t.orth_center
        if start <= end:
            return t
        qr_sweep_right(t.cores, start, end)
    elif canonical == "mix":
        if center is None:
            center = d // 2
        end = center
        qr_sweep_left(t.cores, 0, center)
        qr_sweep_right(t.cores, d - 1, center)
    else:
        raise ValueError(f"Canonical form must be 'left', 'right' or 'mix', but got {canonical}")
    t.orth_center = end if (had or canonical == "mix" or end in (0, d - 1)) else None
    return t


def compress_svd(t: Train, error: float = 0.0, max_rank=None, canonical="auto", spectra=None) -> Train:
    """In-place SVD sweep, direction by centre (``tensor_train.py:789-850``)."""
    if canonical == "auto":
        canonical = "left" if (t.orth_center is None or t.orth_center <= (t.d - 1) / 2) else "right"
    if canonical == "right":
        svd_sweep_right(t.cores, error, max_rank, spectra)
        t.orth_center = 0
    elif canonical == "left":
        svd_sweep_left(t.cores, error, max_rank, spectra)
        t.orth_center = t.d - 1
    else:
        raise ValueError(f"Canonical form must be 'left', 'right' or 'auto', but got {canonical}")
    return t


def sketch_rank(max_rank, options) -> int:
    """Bond of the random sketch of ``method="randomized"`` (``tensor_train.py:122-140``):
    ``max_rank`` plus ``options["oversampling"]`` (default 10)."""
    if max_rank is None:
        raise ValueError("method='randomized' needs max_rank")
    options = {} if options is None else options
    unknown = set(options) - {"oversampling"}
    if unknown:
        raise ValueError(f"Unknown options for method='randomized': {sorted(unknown)}")
    over = options.get("oversampling", 10)
    if not isinstance(over, numbers.Integral) or over < 0:
        raise ValueError(f"oversampling must be a non-negative integer, but got {over!r}")
    return max_rank + int(over)


def randomized_sweep_left(cores: List[np.ndarray], rank: int, rng: np.random.Generator) -> None:
    """Randomize-then-orthogonalize, left form, in place (``utils.py:139-198``).

    From the last core back, W_i = (core_i contracted with W_{i+1}) times a
    standard-normal block of at most ``rank`` columns drawn from ``rng``
    (cores d-1 .. 1 in that order), scaled by its largest magnitude; then
    from the first core on, core_i <- Q of (unfolding_i @ W_{i+1}) and Q^T
    unfolding_i is absorbed into core_{i+1}."""
    if rank < 1:
        raise ValueError(f"rank must be positive, but got {rank}")
    d = len(cores)
    w = [None] * d + [np.ones((1, 1))]
    for i in range(d - 1, 0, -1):
        c = _absorb_right(cores[i], w[i + 1])
        c = c.reshape(c.shape[0], -1)
        bond = min(rank, c.shape[0], c.shape[1])
        g = rng.standard_normal((c.shape[1], bond))
        wi = c @ g
        top = np.abs(wi).max()
        w[i] = wi / top if top > 0 else wi
    for i in range(d - 1):
        s = cores[i].shape
        m = cores[i].reshape(s[0] * s[1] * s[2], s[3])
        q, _ = sla.qr(m @ w[i + 1], mode="economic")
        cores[i] = q.reshape(s[0], s[1], s[2], -1)
        cores[i + 1] = _absorb_left(q.T @ m, cores[i + 1])


def randomized_sweep_right(cores: List[np.ndarray], rank: int, rng: np.random.Generator) -> None:
    """Mirror image of :func:`randomized_sweep_left` (``utils.py:201-227``): the
    left form on the reversed train with the bonds of each core swapped."""
    mirrored = [c.transpose(3, 1, 2, 0) for c in reversed(cores)]
    randomized_sweep_left(mirrored, rank, rng)
    cores[:] = [np.ascontiguousarray(c.transpose(3, 1, 2, 0)) for c in reversed(mirrored)]


def randomized_hadamard(a_cores, b_cores, rank: int, rng: np.random.Generator) -> List[np.ndarray]:
    """Left-orthonormal cores of the element-wise product of two trains from a
    sketch, the product cores never formed (``utils.py:230-333``, HaTT-1).

    The sketch and the sweep are those of :func:`randomized_sweep_left` on the
    product train, whose core i is a_i (x) b_i over the bonds: contractions
    with it are taken one factor at a time.  W_i is (r_i, s_i, l_i)."""
    if rank < 1:
        raise ValueError(f"rank must be positive, but got {rank}")
    d = len(a_cores)
    w = [None] * d + [np.ones((1, 1, 1))]
    for i in range(d - 1, 0, -1):
        a, b = a_cores[i], b_cores[i]
        r, m, n, rp = a.shape
        s, sp = b.shape[0], b.shape[3]
        c = np.einsum("xky,ukv,yvl->xukl", a.reshape(r, m * n, rp), b.reshape(s, m * n, sp),
                      w[i + 1]).reshape(r * s, -1)
        bond = min(rank, c.shape[0], c.shape[1])
        wi = c @ rng.standard_normal((c.shape[1], bond))
        top = np.abs(wi).max()
        if top > 0:
            wi = wi / top
        w[i] = wi.reshape(r, s, bond)
    cores = []
    carry = np.ones((1, 1, 1))
    for i in range(d):
        a, b = a_cores[i], b_cores[i]
        r, m, n, rp = a.shape
        s, sp = b.shape[0], b.shape[3]
        c0 = carry.shape[0]
        mat = np.einsum("cxu,xky,ukv->ckyv", carry, a.reshape(r, m * n, rp),
                        b.reshape(s, m * n, sp)).reshape(c0 * m * n, rp * sp)
        if i == d - 1:
            cores.append(mat.reshape(c0, m, n, -1))
            break
        q, _ = sla.qr(mat @ w[i + 1].reshape(rp * sp, -1), mode="economic")
        cores.append(q.reshape(c0, m, n, -1))
        carry = (q.T @ mat).reshape(q.shape[1], rp, sp)
    return cores


def multiply_rounded(a: Train, b: Train, error=None, max_rank=None, method="direct", seed=None,
                     options=None) -> Train:
    """``a.multiply(b, truncate=True, ...)`` (``tensor_train.py:328-407``)."""
    if method == "randomized":
        out = Train(randomized_hadamard(a.cores, b.cores, sketch_rank(max_rank, options),
                                        np.random.default_rng(seed)), orth_center=a.d - 1)
        return compress_svd(out, 0.0 if error is None else error, max_rank)
    return truncate(tt_multiply(a, b), 0.0 if error is None else error, max_rank)


def truncate(t: Train, error: float = 0.0, max_rank=None, canonical="auto", spectra=None,
             method: str = "direct", seed=None, options=None) -> Train:
    """Copy, orthonormalizing sweep, SVD sweep (``tensor_train.py:852-940``).

    ``method="randomized"`` replaces the QR sweep by the randomized sweep on a
    sketch of bond ``sketch_rank(max_rank, options)`` drawn from
    ``default_rng(seed)``."""
    if canonical not in ("auto", "left", "right"):
        raise ValueError(f"Canonical form must be 'left', 'right' or 'auto', but got {canonical}")
    if method not in ("direct", "randomized"):
        raise ValueError(f"Method must be 'direct' or 'randomized', but got {method!r}")
    if method == "randomized":
        rank = sketch_rank(max_rank, options)
    out = t.copy()
    form = "left" if canonical == "right" else "right"
    if canonical == "auto":
        form = "left" if (out.orth_center is None or out.orth_center >= out.d / 2) else "right"
    if method == "direct":
        orthonormalize(out, form)
    else:
        rng = np.random.default_rng(seed)
        if form == "left":
            randomized_sweep_left(out.cores, rank, rng)
            out.orth_center = out.d - 1
        else:
            randomized_sweep_right(out.cores, rank, rng)
            out.orth_center = 0
    return compress_svd(out, error, max_rank, spectra=spectra)


# --------------------------------------------------------------------------
# TT arithmetic used by the residual check (tensor_train.py:254-540)
# --------------------------------------------------------------------------


def tt_add(a: Train, b: Train) -> Train:
    """Block-placed sum (``tensor_train.py:254-322``); centre unknown."""
    d = a.d
    out = []
    for i in range(d):
        ca, cb = a.cores[i], b.cores[i]
        left = 1 if i == 0 else ca.shape[0] + cb.shape[0]
        right = 1 if i == d - 1 else ca.shape[3] + cb.shape[3]
        c = np.zeros((left, ca.shape[1], ca.shape[2], right))
        c[: ca.shape[0], :, :, : ca.shape[3]] = ca
        c[left - cb.shape[0]:, :, :, right - cb.shape[3]:] = cb
        out.append(c)
    return Train(out)


def tt_multiply(a: Train, b: Train) -> Train:
    """Element-wise product: core (r s, m, n, r' s') with entries
    a[x, m, n, y] b[u, m, n, v] at ((x, u), m, n, (y, v)) (``tensor_train.py:380-395``)."""
    cores = []
    for p, q in zip(a.cores, b.cores):
        c = np.einsum("xmny,umnv->xumnyv", p, q)
        cores.append(c.reshape(p.shape[0] * q.shape[0], p.shape[1], p.shape[2],
                               p.shape[3] * q.shape[3]))
    return Train(cores)


def tt_scale(a: Train, alpha: float) -> Train:
    """Scalar multiple on the centre core, or core 0 (``tensor_train.py:396-401``)."""
    out = a.copy()
    k = out.orth_center if out.orth_center is not None else 0
    out.cores[k] = alpha * out.cores[k]
    return out


def tt_matvec(A: Train, x: Train) -> Train:
    """TTM @ TT cores ``(Ra*rx, n, 1, Ra'*rx')`` (``tensor_train.py:427-499``)."""
    out = []
    for ca, cx in zip(A.cores, x.cores):
        c = np.einsum("aijb,mjkn->amikbn", ca, cx)
        out.append(c.reshape(ca.shape[0] * cx.shape[0], ca.shape[1], cx.shape[2], ca.shape[3] * cx.shape[3]))
    return Train(out)


def tt_norm(t: Train) -> float:
    """Frobenius norm: centre core if known, else right QR sweep (``tensor_train.py:509-540``)."""
    if t.orth_center is not None:
        return float(np.linalg.norm(t.cores[t.orth_center].reshape(-1)))
    c = [x.reshape(x.shape[0], x.shape[1] * x.shape[2], 1, x.shape[3]).copy() for x in t.cores]
    qr_sweep_right(c, len(c) - 1, 0)
    return float(np.linalg.norm(c[0].reshape(-1)))


# --------------------------------------------------------------------------
# QTT conversions  (tensor_train.py:58-91, 942-1135)
# --------------------------------------------------------------------------


def _power_of(mode: int, base: int) -> int:
    """Exponent q with base**q == mode, or -1 (integer search, ``tensor_train.py:975-980``)."""
    q = 0
    while base ** q < mode:
        q += 1
    return q if base ** q == mode else -1


def tt_to_qtt(t: Train, mode_size: int = 2, error: float = 1e-12, max_rank=None) -> Train:
    """Split every core of mode ``mode_size**q`` into ``q`` cores of mode
    ``mode_size`` (``tensor_train.py:942-1073``).

    The train is first right-orthonormalised; each split is a ``gesvd`` of
    the unfolding (left, m, m_c) x (rest), truncated by ``truncated_bond``
    with ``delta = error / sqrt(#SVDs) * ||t||``, the kept ``U`` emitted and
    ``U^T M`` carried on; at a group boundary the carry is pushed into the
    next core by one more truncated SVD.  Returns a left-canonical train with
    its centre at the last core; the input is not modified.
    """
    if mode_size < 2:
        raise ValueError(f"mode_size must be at least 2, but got {mode_size}")
    if max_rank is not None and max_rank <= 0:
        raise ValueError("max_rank must be a positive integer.")
    ttm = any(c > 1 for c in t.col_modes)
    qs = []
    for i, row in enumerate(t.row_modes):
        q = _power_of(row, mode_size)
        if q < 0:
            raise ValueError(f"Row mode {row} not a power of {mode_size}")
        if ttm and mode_size ** q != t.col_modes[i]:
            raise ValueError("Column mode size of TTM not a power of mode size")
        qs.append(q)
    rs, cs = mode_size, (mode_size if ttm else 1)
    n_svd = max(1, sum(max(0, q - 1) for q in qs) + t.d - 1)
    w = orthonormalize(t.copy(), "right")
    delta = error / np.sqrt(n_svd) * tt_norm(w)
    out = []
    for i, q in enumerate(qs):
        core = w.cores[i]
        left, right = core.shape[0], core.shape[-1]
        for j in range(q - 1, 0, -1):
            mat = core.reshape(left, rs, rs ** j, cs, cs ** j, right).transpose(0, 1, 3, 2, 4, 5)
            mat = mat.reshape(left * rs * cs, -1)
            u, s, _ = sla.svd(mat, full_matrices=False, check_finite=False, lapack_driver="gesvd")
            k = truncated_bond(s, delta, max_rank)
            out.append(u[:, :k].reshape(left, rs, cs, k))
            core = u[:, :k].T @ mat
            left = k
        if i < t.d - 1:
            mat = core.reshape(left * rs * cs, right)
            u, s, _ = sla.svd(mat, full_matrices=False, check_finite=False, lapack_driver="gesvd")
            k = truncated_bond(s, delta, max_rank)
            out.append(u[:, :k].reshape(left, rs, cs, k))
            w.cores[i + 1] = _absorb_left(u[:, :k].T @ mat, w.cores[i + 1])
        else:
            out.append(core.reshape(left, rs, cs, right))
    return Train(out, orth_center=len(out) - 1)


def merge_cores(cores: List[np.ndarray]) -> np.ndarray:
    """Consecutive cores as one core (left, prod rows, prod cols, right), rows
    and columns each in the cores' order (``tensor_train.py:58-91``)."""
    left, right = cores[0].shape[0], cores[-1].shape[-1]
    t = cores[0].reshape(-1, cores[0].shape[-1])
    for c in cores[1:]:
        t = (t @ c.reshape(c.shape[0], -1)).reshape(-1, c.shape[-1])
    rows = [c.shape[1] for c in cores]
    cols = [c.shape[2] for c in cores]
    modes = [x for pair in zip(rows, cols) for x in pair]
    kept = [a for a, m in enumerate(modes) if m > 1]     # size-1 modes need no axis
    perm = [1 + i for i, a in enumerate(kept) if a % 2 == 0]
    perm += [1 + i for i, a in enumerate(kept) if a % 2 == 1]
    t = t.reshape([left] + [modes[a] for a in kept] + [right])
    return t.transpose([0] + perm + [len(kept) + 1]).reshape(left, prod(rows), prod(cols), right)


def qtt_groups(t: Train, physical_indexes) -> List[int]:
    """Number of consecutive cores merged for each target index, with the
    reference's dimension checks (``tensor_train.py:1100-1130``)."""
    ttm = any(c > 1 for c in t.col_modes)
    counts, pos = [], 0
    for i, idx in enumerate(physical_indexes):
        row_t, col_t = (idx if ttm else (idx, 1))
        pr = pc = 1
        n = 0
        while n == 0 or pr != row_t or pc != col_t:
            if pos + n >= t.d:
                raise ValueError(f"Not enough QTT cores to match physical index for core {i}")
            pr *= t.row_modes[pos + n]
            pc *= t.col_modes[pos + n]
            n += 1
            if pr > row_t:
                raise ValueError(f"Row mode size of QTT exceeds original row mode size for core {i}")
            if pc > col_t:
                raise ValueError(
                    f"Column mode size of QTT exceeds original column mode size for core {i}")
        counts.append(n)
        pos += n
    if pos != t.d:
        raise ValueError(f"Physical indexes cover only {pos} of {t.d} QTT cores")
    return counts


def qtt_to_tt(t: Train, physical_indexes, error: float = 1e-12, max_rank=None) -> Train:
    """Merge groups of QTT cores back into cores of the given modes, then
    ``truncate(error, max_rank)`` (``tensor_train.py:1075-1135``)."""
    counts = qtt_groups(t, physical_indexes)
    cores, pos = [], 0
    for n in counts:
        cores.append(merge_cores(t.cores[pos:pos + n]))
        pos += n
    return truncate(Train(cores), error=error, max_rank=max_rank)


def relative_residual(A: Train, b: Train, x: Train, b_norm: Optional[float] = None) -> float:
    """``||A x - b|| / ||b||`` in TT arithmetic (``common.py:118-146``)."""
    r = tt_add(tt_matvec(A, x), tt_scale(b, -1.0))
    if b_norm is None:
        b_norm = tt_norm(b)
    rn = tt_norm(r)
    return rn if b_norm == 0.0 else rn / b_norm


# --------------------------------------------------------------------------
# environments and the one-site local system (common.py, one_site.py)
# --------------------------------------------------------------------------


def env_left(prev: np.ndarray, frame_core: np.ndarray, target_core: np.ndarray,
             op_core: Optional[np.ndarray] = None) -> np.ndarray:
    """One step of the left environment recursion (``common.py:149-188``).

    With the operator: ``E'[t',a',f'] = sum E[t,a,f] F[f,i,f'] A[a,i,j,a'] T[t,j,t']``;
    without: ``E'[t',f'] = sum E[t,f] F[f,i,f'] T[t,i,t']``.  ``F`` is the
    frame core (conjugated; real here), ``T`` the target core, both 3-D.
    """
    if op_core is not None:
        s = np.tensordot(prev, frame_core, axes=([2], [0]))          # t a i f'
        s = np.tensordot(s, op_core, axes=([1, 2], [0, 1]))          # t f' j a'
        return np.tensordot(s, target_core, axes=([0, 2], [0, 1])).transpose(2, 1, 0)
    s = np.tensordot(prev, frame_core, axes=([1], [0]))              # t i f'
    return np.tensordot(s, target_core, axes=([0, 1], [0, 1])).T


def env_right(nxt: np.ndarray, frame_core: np.ndarray, target_core: np.ndarray,
              op_core: Optional[np.ndarray] = None) -> np.ndarray:
    """Mirror recursion (``common.py:191-217``), indexed (target, op, frame)."""
    if op_core is not None:
        s = np.tensordot(target_core, nxt, axes=([2], [0]))          # t j a' f'
        s = np.tensordot(s, op_core, axes=([1, 2], [2, 3]))          # t f' a i
        return np.tensordot(s, frame_core, axes=([1, 3], [2, 1]))    # t a f
    s = np.tensordot(target_core, nxt, axes=([2], [0]))              # t i f'
    return np.tensordot(s, frame_core, axes=([1, 2], [1, 2]))


def _c3(core4: np.ndarray) -> np.ndarray:
    return core4[:, :, 0, :]


def build_env(envs: list, k: int, side: str, frame: Train, target: Train, A: Optional[Train] = None):
    """Set ``envs[k]`` from its neighbour (boundary = ones), as the reference does."""
    with_op = A is not None
    if side == "left":
        if k == 0:
            envs[0] = np.ones((1, 1, 1)) if with_op else np.ones((1, 1))
            return
        envs[k] = env_left(envs[k - 1], _c3(frame.cores[k - 1]).conj(), _c3(target.cores[k - 1]),
                           A.cores[k - 1] if with_op else None)
    else:
        if k == frame.d - 1:
            envs[-1] = np.ones((1, 1, 1)) if with_op else np.ones((1, 1))
            return
        envs[k] = env_right(envs[k + 1], _c3(frame.cores[k + 1]).conj(), _c3(target.cores[k + 1]),
                            A.cores[k + 1] if with_op else None)


def local_rhs(psi_b: np.ndarray, b_core3: np.ndarray, phi_b: np.ndarray) -> np.ndarray:
    """``b_loc[f_l,i,f_r] = sum Psi[t,f_l] b[t,i,t'] Phi[t',f_r]`` (``one_site.py:27-46``)."""
    return np.einsum("ai,ajc,cl->ijl", psi_b, b_core3, phi_b, optimize="optimal").reshape(-1)


def local_matrix(psi: np.ndarray, op_core: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Explicit one-site operator, rows (f_l,i,f_r), cols (t_l,j,t_r) (``one_site.py:49-75``)."""
    m = np.einsum("iaj,aklb,mbn->ijklmn", psi, op_core, phi, optimize="optimal")
    n_row, n_col = op_core.shape[1], op_core.shape[2]
    rows = psi.shape[2] * n_row * phi.shape[2]
    cols = psi.shape[0] * n_col * phi.shape[0]
    return m.transpose([1, 2, 5, 0, 3, 4]).reshape(rows, cols)


def local_apply(psi: np.ndarray, op_core: np.ndarray, phi: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Matrix-free one-site apply on columns of ``y`` (``one_site.py:78-115``).

    ``y`` has shape ``(N,)`` or ``(N, B)`` with rows indexed (t_l, j, t_r);
    the result rows are (f_l, i, f_r).
    """
    rl, rr = psi.shape[0], phi.shape[0]
    vec = y.ndim == 1
    yy = y.reshape(rl, op_core.shape[2], rr, -1)
    t = np.tensordot(yy, phi, axes=(2, 0))                      # i l z b n
    t = np.tensordot(t, op_core, axes=((1, 3), (2, 3)))          # i z n a k
    t = np.tensordot(t, psi, axes=((0, 3), (0, 1)))              # z n k j
    out = t.transpose(3, 2, 1, 0).reshape(psi.shape[2] * op_core.shape[1] * phi.shape[2], -1)
    return out.ravel() if vec else out


class _LocalOp(spla.LinearOperator):
    def __init__(self, psi, op_core, phi):
        n = psi.shape[0] * op_core.shape[1] * phi.shape[0]
        super().__init__(dtype=np.float64, shape=(n, n))
        self.psi, self.op_core, self.phi = psi, op_core, phi

    def _matvec(self, v):
        return local_apply(self.psi, self.op_core, self.phi, np.asarray(v).reshape(-1))

    def _matmat(self, m):
        return local_apply(self.psi, self.op_core, self.phi, np.asarray(m))


# constants of common.py:220-241
LOCAL_TOL_FACTOR = 0.1
DEFAULT_LOCAL_RTOL = 1e-8
MAX_LOCAL_RESTARTS = 4
RESTART_RTOL_SHRINK = 1e-2
DEFAULT_STAGNATION = 0.9


@dataclass
class Settings:
    """Resolved options (``common.py:244-400``)."""

    max_sweeps: int
    residual_tol: Optional[float]
    stagnation_factor: Optional[float]
    iterative: Optional[str]
    direct_solve_size: int
    track_residual: bool
    local_rtol: Optional[float]
    allow_dense_fallback: bool = False
    truncation_steps: int = 1

    @property
    def residual_mode(self) -> bool:
        return self.residual_tol is not None

    @property
    def truncation_tol(self) -> Optional[float]:
        return None if self.residual_tol is None else self.residual_tol / np.sqrt(self.truncation_steps)

    def use_iterative(self, n: int) -> bool:
        return self.iterative is not None and n > self.direct_solve_size


def resolve_settings(max_sweeps, residual_tol, stagnation_factor, local_solver, symmetric,
                     direct_solve_size, return_info, local_rtol=None, allow_dense_fallback=False,
                     truncation_steps=1) -> Settings:
    """Validation and derived fields (``common.py:301-400``)."""
    if residual_tol is not None and residual_tol <= 0.0:
        raise ValueError(f"residual_tol must be positive, got {residual_tol}")
    if truncation_steps < 1:
        raise ValueError(f"truncation_steps must be positive, got {truncation_steps}")
    if isinstance(stagnation_factor, str):
        if stagnation_factor != "auto":
            raise ValueError("stagnation_factor must be a float in (0, 1], None or 'auto'")
        stagnation_factor = DEFAULT_STAGNATION if (residual_tol is not None or return_info) else None
    if stagnation_factor is not None and not 0.0 < stagnation_factor <= 1.0:
        raise ValueError(f"stagnation_factor must be in (0, 1], got {stagnation_factor}")
    track = return_info or residual_tol is not None or (stagnation_factor is not None and max_sweeps > 1)
    if local_rtol is not None and local_rtol <= 0.0:
        raise ValueError(f"local_rtol must be positive, got {local_rtol}")
    ls = str(local_solver)
    if ls not in ("auto", "direct", "minres", "gmres"):
        raise ValueError(f"local_solver must be one of 'auto', 'direct', 'minres', 'gmres', got {local_solver!r}")
    iterative = ("minres" if symmetric else "gmres") if ls == "auto" else (None if ls == "direct" else ls)
    if iterative is None:
        lr = None
    elif local_rtol is not None:
        lr = local_rtol
    elif residual_tol is not None:
        lr = LOCAL_TOL_FACTOR * residual_tol
    else:
        lr = DEFAULT_LOCAL_RTOL
    return Settings(max_sweeps, residual_tol, stagnation_factor, iterative, direct_solve_size,
                    track, lr, allow_dense_fallback, truncation_steps)


@dataclass
class Info:
    """Same fields as the reference ``SolverInfo`` (``common.py:85-115``)."""

    residuals: List[float] = field(default_factory=list)
    num_sweeps_run: int = 0
    stop_reason: str = "max_sweeps"
    max_bonds: List[int] = field(default_factory=list)
    local_solves: int = 0
    local_fallbacks: int = 0
    local_inexact: int = 0


def solve_local(op, b_vec, iterative, local_rtol, warm=None, fallback=None, info: Optional[Info] = None):
    """Direct LU or restarted Krylov solve (``common.py:729-809``)."""
    if iterative is None:
        return np.asarray(sla.solve(op, b_vec, check_finite=False)).ravel(), op
    target = local_rtol * np.linalg.norm(b_vec)
    u = warm
    rt = local_rtol
    for _ in range(MAX_LOCAL_RESTARTS):
        if iterative == "minres":
            u, _ = spla.minres(op, b_vec, x0=u, rtol=rt)
        else:
            u, _ = spla.gmres(op, b_vec, x0=u, rtol=rt, atol=0.0)
        if np.linalg.norm(op @ u - b_vec) <= target:
            if info is not None:
                info.local_solves += 1
            return np.asarray(u).ravel(), op
        rt = max(rt * RESTART_RTOL_SHRINK, 1e-15)
    if fallback is None:
        if info is not None:
            info.local_solves += 1
            info.local_inexact += 1
        return np.asarray(u).ravel(), op
    if info is not None:
        info.local_solves += 1
        info.local_fallbacks += 1
    op = fallback()
    return np.asarray(sla.solve(op, b_vec, check_finite=False)).ravel(), op


def rank_by_residual(op, b_loc, U, S, Vt, exact_res, trunc_tol, damp, max_rank,
                     trace: Optional[dict] = None) -> Tuple[int, bool]:
    """Smallest rank whose local residual meets the budget (``common.py:812-866``).

    With ``trace``, records the threshold and the norm of every partial sum
    tested (``thr``, ``partials``), which the site-step parity tests use to
    tell decisions that round-off can flip from those it cannot.
    """
    if S[0] == 0.0:
        if trace is not None:
            trace.update(thr=0.0, partials=[], zero_spectrum=True)
        return 1, False
    limit = len(S) if max_rank is None else min(len(S), max_rank)
    thr = max(damp * exact_res, 0.5 * trunc_tol * np.linalg.norm(b_loc))
    terms = np.einsum("ik,kj->kij", U[:, :limit] * S[:limit], Vt[:limit, :], optimize="optimal")
    contrib = op @ terms.reshape(limit, -1).T
    partial = -b_loc.copy()
    norms = []
    if trace is not None:
        trace.update(thr=thr, partials=norms, exact=exact_res, b_norm=float(np.linalg.norm(b_loc)))
    for r in range(limit):
        partial += contrib[:, r]
        nrm = np.linalg.norm(partial)
        norms.append(float(nrm))
        if nrm <= thr:
            return r + 1, False
    return limit, max_rank is not None and max_rank < len(S)


def rank_by_threshold(S, thr, max_rank) -> Tuple[int, bool]:
    """sigma_i / sigma_1 > thr count (``common.py:872-885``)."""
    if S.size == 0 or S[0] == 0.0:
        return 1, False
    r = max(int(np.count_nonzero(S / S[0] > thr)), 1)
    if max_rank is not None and r > max_rank:
        return max_rank, True
    return r, False


def _choose_rank(settings, op, b_loc, U, S, Vt, exact, error_threshold, max_rank, damp, trace=None):
    if trace is not None:
        trace["S"] = S
    if settings.residual_mode:
        r = rank_by_residual(op, b_loc, U, S, Vt, exact, settings.truncation_tol, damp, max_rank,
                             trace)[0]
    else:
        r = rank_by_threshold(S, error_threshold, max_rank)[0]
    if trace is not None:
        trace["rank"] = r
    return r


class _System:
    """Environments of one sweep solver, held per site (``amen.py:538-566``)."""

    def __init__(self, d):
        self.L = {}
        self.R = {}

    def env(self, side, name, d):
        store = self.L if side == "left" else self.R
        if name not in store:
            store[name] = [None] * d
        return store[name]


def _solve_site(k, x, A, b, eL, eR, bL, bR, settings, info):
    """Assemble and solve the one-site system at ``k`` (``one_site.py:118-178``)."""
    n = x.bonds[k] * x.row_modes[k] * x.bonds[k + 1]
    b_loc = local_rhs(bL[k], _c3(b.cores[k]), bR[k])
    if settings.use_iterative(n):
        op = _LocalOp(eL[k], A.cores[k], eR[k])
        warm = _c3(x.cores[k]).reshape(-1)
        if warm.size != n:
            warm = None
        fb = (lambda: local_matrix(eL[k], A.cores[k], eR[k])) if settings.allow_dense_fallback else None
        u, op = solve_local(op, b_loc, settings.iterative, settings.local_rtol, warm, fb, info)
    else:
        op = local_matrix(eL[k], A.cores[k], eR[k])
        u, op = solve_local(op, b_loc, None, None)
    return u, op, b_loc


def shift_right(x: Train, k: int, core3: np.ndarray) -> None:
    """QR of (r_l n x r_r), Q stored at k, R into k+1 (``one_site.py:181-193``)."""
    q, r = sla.qr(core3.reshape(-1, core3.shape[-1]), mode="economic")
    x.cores[k] = q.reshape(core3.shape[0], core3.shape[1], 1, q.shape[1])
    x.cores[k + 1] = _absorb_left(r, x.cores[k + 1])
    x.orth_center = k + 1


def shift_left(x: Train, k: int, core3: np.ndarray) -> None:
    """QR of the transposed right unfolding, R^T into k-1 (``one_site.py:196-209``)."""
    q, r = sla.qr(core3.reshape(core3.shape[0], -1).T, mode="economic")
    x.cores[k] = q.T.reshape(q.shape[1], core3.shape[1], 1, core3.shape[2])
    x.cores[k - 1] = _absorb_right(x.cores[k - 1], r.T)
    x.orth_center = k - 1


def residual_core(k, zbL, zbR, zAL, zAR, A, b, x_core3):
    """Projection of ``b - A x`` onto z's frame at ``k`` (``amen.py:109-159``)."""
    b_part = local_rhs(zbL[k], _c3(b.cores[k]), zbR[k]).reshape(zbL[k].shape[1], b.cores[k].shape[1], zbR[k].shape[1])
    left, core, right = zAL[k], A.cores[k], zAR[k]
    xl, al, zl = left.shape
    _, rows, cols, ar = core.shape
    xr, _, zr = right.shape
    s = left.transpose(2, 1, 0).reshape(zl * al, xl) @ x_core3.reshape(xl, cols * xr)
    s = s.reshape(zl, al, cols, xr).transpose(0, 3, 1, 2).reshape(zl * xr, al * cols)
    s = s @ core.transpose(0, 2, 1, 3).reshape(al * cols, rows * ar)
    s = s.reshape(zl, xr, rows, ar).transpose(0, 2, 1, 3).reshape(zl * rows, xr * ar) @ right.reshape(xr * ar, zr)
    return b_part - s.reshape(zl, rows, zr)


def fit_residual(A, b, x, z, sweeps=1):
    """ALS fit of z to ``b - A x`` (``amen.py:162-224``)."""
    d = z.d
    zbR, zAR, zbL, zAL = [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, -1, -1):
        build_env(zbR, i, "right", z, b)
        build_env(zAR, i, "right", z, x, A)
    for _ in range(sweeps):
        for k in range(d - 1):
            build_env(zbL, k, "left", z, b)
            build_env(zAL, k, "left", z, x, A)
            shift_right(z, k, residual_core(k, zbL, zbR, zAL, zAR, A, b, _c3(x.cores[k])))
        build_env(zbL, d - 1, "left", z, b)
        build_env(zAL, d - 1, "left", z, x, A)
        for k in range(d - 1, -1, -1):
            build_env(zbR, k, "right", z, b)
            build_env(zAR, k, "right", z, x, A)
            c = residual_core(k, zbL, zbR, zAL, zAR, A, b, _c3(x.cores[k]))
            if k > 0:
                shift_left(z, k, c)
            else:
                z.cores[0] = c.reshape(z.bonds[0], c.shape[1], 1, z.bonds[1])
                z.orth_center = 0
    return z


def _exact_residual(op, u, b_loc):
    return float(np.linalg.norm(op @ u - b_loc))


def truncate_core_right(core3, settings, error_threshold, max_rank, op, b_loc, damp, trace=None):
    """SVD split of (r_l n x r_r) (``amen.py:227-275``): (U_r, S V_r)."""
    rl, n, rr = core3.shape
    exact = _exact_residual(op, core3.ravel(), b_loc) if settings.residual_mode else 0.0
    U, S, Vt = sla.svd(core3.reshape(rl * n, rr), full_matrices=False, check_finite=False, lapack_driver="gesvd")
    r = _choose_rank(settings, op, b_loc, U, S, Vt, exact, error_threshold, max_rank, damp, trace)
    return U[:, :r].reshape(rl, n, r), S[:r, None] * Vt[:r]


def truncate_core_left(core3, settings, error_threshold, max_rank, op, b_loc, damp, trace=None):
    """Mirror (``amen.py:278-322``): (V_r as (r,n,r_r), U_r S)."""
    rl, n, rr = core3.shape
    exact = _exact_residual(op, core3.ravel(), b_loc) if settings.residual_mode else 0.0
    U, S, Vt = sla.svd(core3.reshape(rl, n * rr), full_matrices=False, check_finite=False, lapack_driver="gesvd")
    r = _choose_rank(settings, op, b_loc, U, S, Vt, exact, error_threshold, max_rank, damp, trace)
    return Vt[:r].reshape(r, n, rr), U[:, :r] * S[:r]


def enrich_right(x, k, core, carry, block, max_rank):
    """Append ``block`` on the right bond, zero-padded carry into k+1 (``amen.py:325-360``)."""
    enriched = np.concatenate([core, block], axis=2)
    pad = np.zeros((block.shape[2], carry.shape[1]))
    x.cores[k + 1] = _absorb_left(np.concatenate([carry, pad], axis=0), x.cores[k + 1])
    shift_right(x, k, enriched)
    if max_rank is not None and x.bonds[k + 1] > max_rank:
        x.cores[k] = x.cores[k][:, :, :, :max_rank]
        x.cores[k + 1] = x.cores[k + 1][:max_rank]


def enrich_left(x, k, core, carry, block, max_rank):
    """Mirror (``amen.py:363-392``)."""
    enriched = np.concatenate([core, block], axis=0)
    pad = np.zeros((carry.shape[0], block.shape[0]))
    x.cores[k - 1] = _absorb_right(x.cores[k - 1], np.concatenate([carry, pad], axis=1))
    shift_left(x, k, enriched)
    if max_rank is not None and x.bonds[k] > max_rank:
        x.cores[k] = x.cores[k][:max_rank]
        x.cores[k - 1] = x.cores[k - 1][:, :, :, :max_rank]


def _record(settings, info, A, b, x, b_norm, sweep, max_bond=None) -> bool:
    """Per-sweep residual and stop test (``common.py:504-592``)."""
    info.num_sweeps_run = sweep + 1
    if max_bond is not None:
        info.max_bonds.append(max_bond)
    if not settings.track_residual:
        return False
    res = relative_residual(A, b, x, b_norm)
    info.residuals.append(res)
    if settings.residual_tol is not None and res <= settings.residual_tol:
        info.stop_reason = "converged"
        return True
    if (settings.stagnation_factor is not None and len(info.residuals) >= 2
            and info.residuals[-1] > settings.stagnation_factor * info.residuals[-2]):
        info.stop_reason = "stagnated"
        return True
    return False


def _validate(A: Train, b: Train, x0: Train):
    """Structural checks (``common.py:684-726``)."""
    if any(m != 1 for m in x0.col_modes):
        raise TypeError("Initial guess x0 must be a TensorTrain, not a TTM.")
    if any(m != 1 for m in b.col_modes):
        raise TypeError("Right-hand side b must be a TensorTrain, not a TTM.")
    if not (A.d == b.d == x0.d):
        raise ValueError("A, b and x0 must have the same number of cores")
    if A.row_modes != A.col_modes or A.row_modes != b.row_modes or A.col_modes != x0.row_modes:
        raise ValueError("dimension mismatch")


@dataclass
class SiteState:
    """Snapshot of the AMEn state before one site step (``amen.py:579-708``).

    ``x`` and ``z`` are the cores of the solution and residual trains at
    that moment (shared arrays: the solver rebinds cores, it never writes
    into them), ``envs`` the ten environments at site ``k`` keyed
    ``A_L``/``A_R`` (x,x,A), ``b_L``/``b_R`` (x,b), ``zb_L``/``zb_R`` (z,b),
    ``zA_L``/``zA_R`` (z,x,A) and ``xz_L``/``xz_R`` (x,z).
    """

    sweep: int
    direction: str
    k: int
    x: List[np.ndarray]
    x_center: Optional[int]
    z: List[np.ndarray]
    envs: dict


def site_step(state: SiteState, A: Train, b: Train, settings: Settings, error_threshold=1e-7,
              max_rank=None, residual_damp=2.0) -> dict:
    """Re-run one AMEn site step from a snapshot, as ``amen`` runs it.

    Returns the step's intermediate results: ``b_loc``, the local solution
    ``u``, the residual core ``z_new`` (with its two terms' norm
    ``z_scale``), the truncation ``trace`` (spectrum, threshold, partial
    norms, rank), ``core``/``carry``/``block`` and the cores of x and z at
    ``k`` and its neighbour after the step.  Nothing in ``state`` changes.
    """
    k, d = state.k, A.d
    x = Train(list(state.x))
    x.orth_center = state.x_center
    z = Train(list(state.z))
    e = state.envs

    def at(name):
        lst = [None] * d
        lst[k] = e[name]
        return lst

    out = {}
    u, op, b_loc = _solve_site(k, x, A, b, at("A_L"), at("A_R"), at("b_L"), at("b_R"), settings, Info())
    core3 = u.reshape(x.bonds[k], x.row_modes[k], x.bonds[k + 1])
    z_new = residual_core(k, at("zb_L"), at("zb_R"), at("zA_L"), at("zA_R"), A, b, core3)
    b_part = local_rhs(e["zb_L"], _c3(b.cores[k]), e["zb_R"])
    out.update(b_loc=b_loc, u=u, z_new=z_new,
               z_scale=float(np.linalg.norm(b_part) + np.linalg.norm(z_new.ravel() - b_part)))
    trace = {}
    if state.direction == "forward":
        core, carry = truncate_core_right(core3, settings, error_threshold, max_rank, op, b_loc,
                                          residual_damp, trace)
        block = np.einsum("ba,bin->ain", e["xz_L"], z_new)
        enrich_right(x, k, core, carry, block, max_rank)
        shift_right(z, k, z_new)
        nb = k + 1
    elif k > 0:
        core, carry = truncate_core_left(core3, settings, error_threshold, max_rank, op, b_loc,
                                         residual_damp, trace)
        block = _absorb_right(z_new, e["xz_R"])
        enrich_left(x, k, core, carry, block, max_rank)
        shift_left(z, k, z_new)
        nb = k - 1
    else:
        return out
    out.update(trace=trace, core=core, carry=carry, block=block, neighbour=nb,
               x_k=x.cores[k], x_nb=x.cores[nb], z_k=z.cores[k], z_nb=z.cores[nb])
    return out


def amen(A: Train, b: Train, x0: Optional[Train] = None, kickrank=4, error_threshold=1e-7,
         max_rank=None, max_sweeps=1, residual_tol=None, residual_damp=2.0, stagnation_factor="auto",
         local_solver="auto", direct_solve_size=1000, local_rtol=None, allow_dense_fallback=False,
         symmetric=False, seed=None, return_info=True, hook=None):
    """One-site AMEn (``amen.py:395-725``).  Returns ``(x, Info)``.

    ``hook``, when given, is called before every site step with a
    :class:`SiteState` snapshot of the full solver state at that step (x, z,
    and every environment the step reads); it changes nothing in the run.
    """
    if kickrank <= 0:
        raise ValueError(f"kickrank must be a positive integer, got {kickrank}")
    if max_rank is not None and max_rank <= 0:
        raise ValueError(f"max_rank must be a positive integer, got {max_rank}")
    if residual_damp < 1.0:
        raise ValueError(f"residual_damp must be >= 1, got {residual_damp}")
    st = resolve_settings(max_sweeps, residual_tol, stagnation_factor, local_solver, symmetric,
                          direct_solve_size, return_info, local_rtol, allow_dense_fallback,
                          max(2 * (A.d - 1), 1))
    if x0 is None:
        x0 = Train([np.ones((b.bonds[i], b.row_modes[i], 1, b.bonds[i + 1])) for i in range(b.d)])
    _validate(A, b, x0)
    x = x0.copy()
    if x.orth_center != 0:
        orthonormalize(x, "right")
    d = x.d
    eA_L, eA_R, eb_L, eb_R = [None] * d, [None] * d, [None] * d, [None] * d
    rng = np.random.default_rng(seed)
    z = Train([rng.random((([1] + [kickrank] * (d - 1) + [1])[i], b.row_modes[i], 1,
                           ([1] + [kickrank] * (d - 1) + [1])[i + 1])) for i in range(d)])
    orthonormalize(z, "right")
    fit_residual(A, b, x, z, 1)
    zbL, zbR, zAL, zAR, xzL, xzR = [None] * d, [None] * d, [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, -1, -1):
        build_env(eA_R, i, "right", x, x, A)
        build_env(eb_R, i, "right", x, b)
        build_env(zbR, i, "right", z, b)
        build_env(zAR, i, "right", z, x, A)
        build_env(xzR, i, "right", x, z)
    info = Info()
    b_norm = tt_norm(b) if st.track_residual else None

    def left_envs(k):
        build_env(eA_L, k, "left", x, x, A)
        build_env(eb_L, k, "left", x, b)
        build_env(zbL, k, "left", z, b)
        build_env(zAL, k, "left", z, x, A)
        build_env(xzL, k, "left", x, z)

    def right_envs(k):
        build_env(eA_R, k, "right", x, x, A)
        build_env(eb_R, k, "right", x, b)
        build_env(zbR, k, "right", z, b)
        build_env(zAR, k, "right", z, x, A)
        build_env(xzR, k, "right", x, z)

    def snap(sweep, direction, k):
        if hook is not None:
            hook(SiteState(sweep, direction, k, list(x.cores), x.orth_center, list(z.cores),
                           {n: e[k] for n, e in (("A_L", eA_L), ("A_R", eA_R), ("b_L", eb_L),
                                                 ("b_R", eb_R), ("zb_L", zbL), ("zb_R", zbR),
                                                 ("zA_L", zAL), ("zA_R", zAR), ("xz_L", xzL),
                                                 ("xz_R", xzR))}))

    for sweep in range(st.max_sweeps):
        for k in range(d - 1):
            left_envs(k)
            snap(sweep, "forward", k)
            u, op, b_loc = _solve_site(k, x, A, b, eA_L, eA_R, eb_L, eb_R, st, info)
            core3 = u.reshape(x.bonds[k], x.row_modes[k], x.bonds[k + 1])
            z_new = residual_core(k, zbL, zbR, zAL, zAR, A, b, core3)
            core, carry = truncate_core_right(core3, st, error_threshold, max_rank, op, b_loc, residual_damp)
            block = np.einsum("ba,bin->ain", xzL[k], z_new)
            enrich_right(x, k, core, carry, block, max_rank)
            shift_right(z, k, z_new)
        left_envs(d - 1)
        for k in range(d - 1, -1, -1):
            right_envs(k)
            snap(sweep, "backward", k)
            u, op, b_loc = _solve_site(k, x, A, b, eA_L, eA_R, eb_L, eb_R, st, info)
            core3 = u.reshape(x.bonds[k], x.row_modes[k], x.bonds[k + 1])
            z_new = residual_core(k, zbL, zbR, zAL, zAR, A, b, core3)
            if k == 0:
                x.cores[0] = core3.reshape(x.bonds[0], core3.shape[1], 1, x.bonds[1])
                x.orth_center = 0
                z.cores[0] = z_new.reshape(z.bonds[0], z_new.shape[1], 1, z.bonds[1])
                z.orth_center = 0
            else:
                core, carry = truncate_core_left(core3, st, error_threshold, max_rank, op, b_loc, residual_damp)
                block = _absorb_right(z_new, xzR[k])
                enrich_left(x, k, core, carry, block, max_rank)
                shift_left(z, k, z_new)
        if _record(st, info, A, b, x, b_norm, sweep, max(x.bonds)):
            break
    redundant = [k for k in range(1, d) if x.bonds[k] > prod(x.row_modes[:k])]
    if redundant:
        orthonormalize(x, "left", redundant[-1])
        orthonormalize(x, "right", 0)
    return x, info


def _merge_pair(c0: np.ndarray, c1: np.ndarray) -> np.ndarray:
    """Two neighbouring cores (r, m0, n0, s), (s, m1, n1, t) as one core
    (r, m0 m1, n0 n1, t): a two-site system is the one-site system of this
    core (``two_site.py:28-119`` index the rows (bond, m0, m1, bond) and the
    columns (bond, n0, n1, bond) the same way)."""
    r, m0, n0, _ = c0.shape
    _, m1, n1, t = c1.shape
    both = np.tensordot(c0, c1, axes=([3], [0]))          # r m0 n0 m1 n1 t
    return both.transpose(0, 1, 3, 2, 4, 5).reshape(r, m0 * m1, n0 * n1, t)


def _solve_pair(k, x, bonds, A, b, eL, eR, bL, bR, settings, info, center, error_threshold,
                max_rank, damp):
    """Assemble, solve and SVD-split the two-site system at (k, k+1)
    (``two_site.py:147-241``).  ``bonds`` are the current bonds (the cores at
    k+1 may be stale mid-sweep); ``center`` = (site, block) replaces the stale
    core in the warm start (``two_site.py:122-144``)."""
    rows = bonds[k] * x.row_modes[k]
    cols = x.row_modes[k + 1] * bonds[k + 2]
    n = rows * cols
    Ak = _merge_pair(A.cores[k], A.cores[k + 1])
    b_loc = local_rhs(bL[k], _c3(_merge_pair(b.cores[k], b.cores[k + 1])), bR[k + 1])
    residual_mode = settings.residual_tol is not None
    if settings.use_iterative(n):
        op = _LocalOp(eL[k], Ak, eR[k + 1])
        left, right = _c3(x.cores[k]), _c3(x.cores[k + 1])
        if center is not None:
            if center[0] == k:
                left = center[1]
            elif center[0] == k + 1:
                right = center[1]
        warm = None
        if left.shape[-1] == right.shape[0]:
            warm = np.tensordot(left, right, axes=(2, 0)).reshape(-1)
            if warm.size != n:
                warm = None
        fb = (lambda: local_matrix(eL[k], Ak, eR[k + 1])) if settings.allow_dense_fallback else None
        u, op = solve_local(op, b_loc, settings.iterative, settings.local_rtol, warm, fb, info)
    else:
        op = local_matrix(eL[k], Ak, eR[k + 1])
        u, op = solve_local(op, b_loc, None, None)
    exact = float(np.linalg.norm(op @ u - b_loc)) if residual_mode else 0.0
    U, S, Vt = sla.svd(u.reshape(rows, cols), full_matrices=False, check_finite=False,
                       lapack_driver="gesvd")
    nb = _choose_rank(settings, op, b_loc, U, S, Vt, exact, error_threshold, max_rank, damp)
    return U, S, Vt, nb


def mals(A: Train, b: Train, x0: Optional[Train] = None, max_rank=None, max_sweeps=2,
         residual_tol=None, residual_damp=2.0, stagnation_factor="auto", error_threshold=1e-7,
         local_solver="auto", direct_solve_size=1000, local_rtol=None, allow_dense_fallback=False,
         symmetric=False, return_info=True):
    """Two-site MALS (``mals.py:61-292``).  Returns ``(x, Info)``.

    Left-to-right over pairs (k, k+1), k < d-2: core k <- U_r, the block
    S_r V_r kept as the centre; right-to-left over all pairs: core k+1 <- V_r,
    the block U_r S_r kept, and at k = 0 stored as core 0.  Truncations per
    sweep: 2d - 3 (the residual budget is shared among them)."""
    if max_rank is not None and max_rank <= 0:
        raise ValueError(f"max_rank must be a positive integer, got {max_rank}")
    if residual_damp < 1.0:
        raise ValueError(f"residual_damp must be >= 1, got {residual_damp}")
    st = resolve_settings(max_sweeps, residual_tol, stagnation_factor, local_solver, symmetric,
                          direct_solve_size, return_info, local_rtol, allow_dense_fallback,
                          max(2 * A.d - 3, 1))
    if x0 is None:
        x0 = Train([np.ones((b.bonds[i], b.row_modes[i], 1, b.bonds[i + 1])) for i in range(b.d)])
    _validate(A, b, x0)
    x = x0.copy()
    if x.orth_center != 0:
        orthonormalize(x, "right")
    d = x.d
    bonds = list(x.bonds)
    eA_L, eA_R, eb_L, eb_R = [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, 0, -1):
        build_env(eA_R, i, "right", x, x, A)
        build_env(eb_R, i, "right", x, b)
    info = Info()
    b_norm = tt_norm(b) if st.track_residual else None
    center = None
    for sweep in range(st.max_sweeps):
        for k in range(d - 1):
            build_env(eA_L, k, "left", x, x, A)
            build_env(eb_L, k, "left", x, b)
            if k == d - 2:
                continue
            U, S, Vt, nb = _solve_pair(k, x, bonds, A, b, eA_L, eA_R, eb_L, eb_R, st, info, center,
                                       error_threshold, max_rank, residual_damp)
            n1 = x.row_modes[k + 1]
            center = (k + 1, (S[:nb, None] * Vt[:nb]).reshape(nb, n1, bonds[k + 2]))
            x.cores[k] = U[:, :nb].reshape(bonds[k], x.row_modes[k], 1, nb)
            bonds[k + 1] = nb
            x.orth_center = k + 1
        for k in range(d - 2, -1, -1):
            build_env(eA_R, k + 1, "right", x, x, A)
            build_env(eb_R, k + 1, "right", x, b)
            U, S, Vt, nb = _solve_pair(k, x, bonds, A, b, eA_L, eA_R, eb_L, eb_R, st, info, center,
                                       error_threshold, max_rank, residual_damp)
            n0, n1 = x.row_modes[k], x.row_modes[k + 1]
            us = U[:, :nb] * S[:nb]
            center = (k, us.reshape(bonds[k], n0, nb))
            x.cores[k + 1] = Vt[:nb].reshape(nb, n1, 1, bonds[k + 2])
            bonds[k + 1] = nb
            if k == 0:
                x.cores[0] = us.reshape(bonds[0], n0, 1, nb)
            x.orth_center = k
        if _record(st, info, A, b, x, b_norm, sweep, max_bond=max(x.bonds)):
            break
    return x, info


def als(A: Train, b: Train, x0: Optional[Train] = None, max_sweeps=4, residual_tol=None,
        stagnation_factor="auto", local_solver="auto", direct_solve_size=1000, local_rtol=None,
        allow_dense_fallback=False, symmetric=False, return_info=True):
    """Fixed-rank one-site ALS (``als.py:56-212``).  Returns ``(x, Info)``."""
    st = resolve_settings(max_sweeps, residual_tol, stagnation_factor, local_solver, symmetric,
                          direct_solve_size, return_info, local_rtol, allow_dense_fallback)
    if x0 is None:
        x0 = Train([np.ones((b.bonds[i], b.row_modes[i], 1, b.bonds[i + 1])) for i in range(b.d)])
    _validate(A, b, x0)
    x = x0.copy()
    if x.orth_center != 0:
        orthonormalize(x, "right")
    d = x.d
    eA_L, eA_R, eb_L, eb_R = [None] * d, [None] * d, [None] * d, [None] * d
    for i in range(d - 1, -1, -1):
        build_env(eA_R, i, "right", x, x, A)
        build_env(eb_R, i, "right", x, b)
    info = Info()
    b_norm = tt_norm(b) if st.track_residual else None
    for sweep in range(st.max_sweeps):
        for k in range(d - 1):
            build_env(eA_L, k, "left", x, x, A)
            build_env(eb_L, k, "left", x, b)
            u, _, _ = _solve_site(k, x, A, b, eA_L, eA_R, eb_L, eb_R, st, info)
            shift_right(x, k, u.reshape(x.bonds[k], x.row_modes[k], x.bonds[k + 1]))
        build_env(eA_L, d - 1, "left", x, x, A)
        build_env(eb_L, d - 1, "left", x, b)
        for k in range(d - 1, -1, -1):
            build_env(eA_R, k, "right", x, x, A)
            build_env(eb_R, k, "right", x, b)
            u, _, _ = _solve_site(k, x, A, b, eA_L, eA_R, eb_L, eb_R, st, info)
            if k == 0:
                x.cores[0] = u.reshape(x.bonds[0], x.row_modes[0], 1, x.bonds[1])
                x.orth_center = 0
            else:
                shift_left(x, k, u.reshape(x.bonds[k], x.row_modes[k], x.bonds[k + 1]))
        if _record(st, info, A, b, x, b_norm, sweep):
            break
    return x, info
