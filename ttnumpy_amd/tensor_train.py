"""Device-resident tensor train with the reference container's interface.

Mirrors the parts of ``ttnumpy.tensor_train.TensorTrain`` (ttnumpy 0.3.0,
``ttnumpy/tensor_train.py:150-940``) on the solver hot path: layout and
validation, ``copy``, ``orthonormalize``, ``compress_svd``, ``truncate``,
``add``, ``multiply`` (scalar and element-wise), ``dot`` (TTM @ TT), ``tt_norm``
and the QTT conversions ``tt_to_qtt`` / ``qtt_to_tt``, plus the
``rand``/``ones``/``zeros``/``eye`` constructors.  Cores are contiguous FP64
device tensors ``(r, n_row, n_col, r')``; every sweep runs on libttx kernels.
"""

from __future__ import annotations

import math
import numbers
import warnings
from enum import Enum
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from . import device as D
from .device import P


class DimensionMismatch(Exception):
    """Exception for dimension errors (``tensor_train.py:21-26``)."""

    def __init__(self, msg="Dimension mismatch"):
        self.msg = msg
        super().__init__(msg)


class WrongTTFormat(Exception):
    """Exception for errors in tensor train structure (``tensor_train.py:29-34``)."""

    def __init__(self, msg="Wrong tensor train format"):
        self.msg = msg
        super().__init__(msg)


class WrongTTType(Exception):
    """Exception for errors in tensor train type (``tensor_train.py:37-42``)."""

    def __init__(self, msg="Wrong tensor train type"):
        self.msg = msg
        super().__init__(msg)


class TensorTrainType(Enum):
    TT = "Tensor Train"
    TTM = "Tensor Train Matrix"


def _as_core(c) -> torch.Tensor:
    if isinstance(c, torch.Tensor):
        t = c.to(device=D.device(), dtype=D.F64)
    else:
        t = D.to_device(np.asarray(c, dtype=np.float64))
    return t.contiguous()


# ---------------------------------------------------------------- sweeps
def qr_sweep_left(cores: List[torch.Tensor], start: int, end: int) -> None:
    """Left QR sweep ``start..end-1`` (``utils.py:86-112``)."""
    for i in range(start, end):
        c = cores[i]
        r0, nr, nc, r1 = c.shape
        m = r0 * nr * nc
        k = min(m, r1)
        Q = D.empty(r0, nr, nc, k)
        R = D.empty(k, r1)
        D.qr(m, r1, P(c), r1, 1, P(Q), k, 1, P(R), r1, 1)
        cores[i] = Q
        cores[i + 1] = D.absorb_left(P(R), k, r1, cores[i + 1])


_WIDE_CARRY_MIN = 1 << 20


def qr_sweep_right(cores: List[torch.Tensor], start: int, end: int, form_q: bool = True) -> None:
    """Right QR sweep ``start..end+1`` on the transposed unfoldings (``utils.py:115-136``).

    With ``form_q=False`` only the carried factors matter (the norm of the
    train ends up in ``cores[end]``).  A transposed unfolding with no more rows
    than columns (w <= r0) has a square orthogonal Q, so its R is the unfolding
    up to a rotation that changes no norm: for such unfoldings of at least
    ``_WIDE_CARRY_MIN`` elements (one-workgroup QRs of milliseconds) the
    unfolding itself is carried instead, by one GEMM.  Smaller ones keep the
    QR, which the batched engine's residual check runs too, so both paths
    compute the same bits."""
    for i in range(start, end, -1):
        c = cores[i]
        r0, nr, nc, r1 = c.shape
        w = nr * nc * r1
        k = min(w, r0)
        if not form_q and w <= r0 and w * r0 >= _WIDE_CARRY_MIN:
            cores[i - 1] = D.absorb_right(cores[i - 1], P(c), r0, w, m_rs=w, m_cs=1)
            continue
        R = D.empty(k, r0)
        if form_q:
            Qt = D.empty(k, nr, nc, r1)
            D.qr(w, r0, P(c), 1, w, P(Qt), 1, w, P(R), r0, 1)
            cores[i] = Qt
        else:
            D.qr(w, r0, P(c), 1, w, None, 1, w, P(R), r0, 1)
        cores[i - 1] = D.absorb_right(cores[i - 1], P(R), r0, k, m_rs=1, m_cs=r0)


def _sweep_status(buf: torch.Tensor, core: int) -> int:
    info, bond = D.sync_ints(buf)
    if info == _lib.TTX_ENONFINITE:
        raise ValueError(f"Core {core} has NaN or infinite entries")
    if info == _lib.TTX_ENOCONV:
        warnings.warn(f"SVD of core {core} did not converge within the sweep limit", RuntimeWarning)
    return bond


def svd_sweep_left(cores: List[torch.Tensor], error: float, max_rank=None, spectra=None) -> None:
    """Left-to-right truncated SVD sweep (``utils.py:364-411``)."""
    d = len(cores)
    nrm = D.norm(cores[0])
    scale = error / math.sqrt(d - 1)
    bond = 1
    buf = torch.zeros(2, dtype=torch.int32, device=D.device())
    for i in range(d - 1):
        c = cores[i]
        _, nr, nc, r1 = c.shape
        m = bond * nr * nc
        U, S, Vt = D.svd(m, r1, P(c), r1, 1, buf, 0)
        k = min(m, r1)
        D.truncated_bond_dev(S, nrm, scale, max_rank, buf, 1)
        nb = _sweep_status(buf, i)
        if spectra is not None:
            spectra.append(S.clone())
        new = D.empty(bond, nr, nc, nb)
        D.copy_nd([m, nb], P(U), [k, 1], P(new), [nb, 1])
        cores[i] = new
        D.scale_cols(r1, nb, P(Vt), 1, r1, P(S))
        cores[i + 1] = D.absorb_left(P(Vt), nb, r1, cores[i + 1])
        bond = nb


def svd_sweep_right(cores: List[torch.Tensor], error: float, max_rank=None, spectra=None) -> None:
    """Right-to-left truncated SVD sweep (``utils.py:414-460``)."""
    d = len(cores)
    nrm = D.norm(cores[-1])
    scale = error / math.sqrt(d - 1)
    bond = 1
    buf = torch.zeros(2, dtype=torch.int32, device=D.device())
    for i in range(d - 1, 0, -1):
        c = cores[i]
        r0, nr, nc, _ = c.shape
        w = nr * nc * bond
        U, S, Vt = D.svd(r0, w, P(c), w, 1, buf, 0)
        k = min(r0, w)
        D.truncated_bond_dev(S, nrm, scale, max_rank, buf, 1)
        nb = _sweep_status(buf, i)
        if spectra is not None:
            spectra.append(S.clone())
        cores[i] = Vt[:nb].reshape(nb, nr, nc, bond)
        D.scale_cols(r0, nb, P(U), k, 1, P(S))
        cores[i - 1] = D.absorb_right(cores[i - 1], P(U), r0, nb, m_rs=k, m_cs=1)
        bond = nb


def _sketch_rank(max_rank, options) -> int:
    """Bond of the random sketch of ``method="randomized"``: ``max_rank`` plus
    ``options["oversampling"]`` (default 10) (``tensor_train.py:122-140``)."""
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


def randomized_sweep_left(cores: List[torch.Tensor], rank: int, rng: np.random.Generator) -> None:
    """Randomize-then-orthogonalize, left form, in place (``utils.py:139-198``).

    From the last core back, W_i = (core_i contracted with W_{i+1}) times a
    standard-normal block of at most ``rank`` columns, scaled by its largest
    magnitude; then from the first core on, core_i becomes the Q factor of
    unfolding_i W_{i+1} and Q^T unfolding_i is absorbed into core_{i+1}.  The
    normal blocks are drawn on the host from ``rng`` (cores d-1 .. 1, the
    reference's generator and order) and copied to the device; the products
    run on the MFMA GEMM and the QRs on ``ttx_qr_f64``.
    """
    if rank < 1:
        raise ValueError(f"rank must be positive, but got {rank}")
    d = len(cores)
    L = D.lib()
    w: List[Optional[torch.Tensor]] = [None] * d + [D.to_device(np.ones((1, 1)))]
    for i in range(d - 1, 0, -1):
        c = cores[i]
        r0, nr, nc, r1 = c.shape
        l1 = w[i + 1].shape[1]
        cols = nr * nc * l1
        contracted = D.empty(r0 * nr * nc, l1)          # = (r0, cols) row-major
        D.gemm(1, r0 * nr * nc, l1, r1, 1.0, P(c), r1, 1, 0, P(w[i + 1]), l1, 1, 0, 0.0,
               P(contracted), l1, 1, 0)
        bond = min(rank, r0, cols)
        g = D.to_device(rng.standard_normal((cols, bond)))
        wi = D.empty(r0, bond)
        D.gemm(1, r0, bond, cols, 1.0, P(contracted), cols, 1, 0, P(g), bond, 1, 0, 0.0, P(wi),
               bond, 1, 0)
        _lib.check(L.ttx_scale_maxabs_f64(D.stream(), r0 * bond, P(wi)), "ttx_scale_maxabs_f64")
        w[i] = wi
    for i in range(d - 1):
        c = cores[i]
        r0, nr, nc, r1 = c.shape
        m = r0 * nr * nc
        l1 = w[i + 1].shape[1]
        mw = D.empty(m, l1)
        D.gemm(1, m, l1, r1, 1.0, P(c), r1, 1, 0, P(w[i + 1]), l1, 1, 0, 0.0, P(mw), l1, 1, 0)
        k = min(m, l1)
        Q = D.empty(r0, nr, nc, k)
        R = D.empty(k, l1)
        D.qr(m, l1, P(mw), l1, 1, P(Q), k, 1, P(R), l1, 1)
        qtm = D.empty(k, r1)                             # Q^T unfolding_i
        D.gemm(1, k, r1, m, 1.0, P(Q), 1, k, 0, P(c), r1, 1, 0, 0.0, P(qtm), r1, 1, 0)
        cores[i] = Q
        cores[i + 1] = D.absorb_left(P(qtm), k, r1, cores[i + 1])


def _product_times_sketch(a: torch.Tensor, b: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """C[(x, u), (k, l)] = sum_{y, v} a[x, k, y] b[u, k, v] w[y, v, l] for cores a
    (r, m, n, r'), b (s, m, n, s') (k = (m, n)) and w (r', s', l): the product
    core of the two trains contracted with the sketch without forming it."""
    r, m, n, rp = a.shape
    s, sp = b.shape[0], b.shape[3]
    mn, l = m * n, w.shape[2]
    h = D.empty(r * mn, sp * l)                      # h[x, k, v, l] = sum_y a w
    D.gemm(1, r * mn, sp * l, rp, 1.0, P(a), rp, 1, 0, P(w), sp * l, 1, 0, 0.0, P(h), sp * l, 1, 0)
    hp = D.empty(mn, sp, r, l)                       # hp[k, v, x, l]
    D.copy_nd([r, mn, sp, l], P(h), [mn * sp * l, sp * l, l, 1], P(hp),
              [l, sp * r * l, r * l, 1])
    c = D.empty(mn, s, r * l)                        # per k: b[:, k, :] hp[k]
    D.gemm(mn, s, r * l, sp, 1.0, P(b), mn * sp, 1, sp, P(hp), r * l, 1, sp * r * l, 0.0, P(c),
           r * l, 1, s * r * l)
    out = D.empty(r * s, mn * l)                     # out[(x, u), (k, l)] = c[k, u, x, l]
    D.copy_nd([mn, s, r, l], P(c), [s * r * l, r * l, l, 1], P(out),
              [l, mn * l, s * mn * l, 1])
    return out


def _carry_times_product(carry: torch.Tensor, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """M[(c, k), (y, v)] = sum_{x, u} carry[c, x, u] a[x, k, y] b[u, k, v]: the carry
    (c, r, s) contracted with the product core of a (r, m, n, r') and b (s, m, n, s')."""
    r, m, n, rp = a.shape
    s, sp = b.shape[0], b.shape[3]
    mn, c0 = m * n, carry.shape[0]
    cp = D.empty(c0, s, r)                           # cp[c, u, x]
    D.copy_nd([c0, r, s], P(carry), [r * s, s, 1], P(cp), [s * r, 1, r])
    t = D.empty(c0 * s, mn * rp)                     # t[c, u, k, y] = sum_x cp a
    D.gemm(1, c0 * s, mn * rp, r, 1.0, P(cp), r, 1, 0, P(a), mn * rp, 1, 0, 0.0, P(t), mn * rp, 1, 0)
    tp = D.empty(mn, c0 * rp, s)                     # tp[k, (c, y), u]
    D.copy_nd([c0, s, mn, rp], P(t), [s * mn * rp, mn * rp, rp, 1], P(tp),
              [rp * s, 1, c0 * rp * s, s])
    cm = D.empty(mn, c0 * rp, sp)                    # per k: tp[k] b[:, k, :]
    D.gemm(mn, c0 * rp, sp, s, 1.0, P(tp), s, 1, c0 * rp * s, P(b), mn * sp, 1, sp, 0.0, P(cm),
           sp, 1, c0 * rp * sp)
    out = D.empty(c0 * mn, rp * sp)                  # out[(c, k), (y, v)] = cm[k, c, y, v]
    D.copy_nd([mn, c0, rp, sp], P(cm), [c0 * rp * sp, rp * sp, sp, 1], P(out),
              [rp * sp, mn * rp * sp, sp, 1])
    return out


def randomized_hadamard(a_cores: List[torch.Tensor], b_cores: List[torch.Tensor], rank: int,
                        rng: np.random.Generator) -> List[torch.Tensor]:
    """Left-orthonormal cores of the element-wise product of two trains from a
    sketch, the product cores never formed (``utils.py:230-333``, HaTT-1): the
    sketch and sweep of :func:`randomized_sweep_left` on the product train,
    each contraction with a product core taken one factor at a time."""
    if rank < 1:
        raise ValueError(f"rank must be positive, but got {rank}")
    d = len(a_cores)
    L = D.lib()
    w: List[Optional[torch.Tensor]] = [None] * d + [D.to_device(np.ones((1, 1, 1)))]
    for i in range(d - 1, 0, -1):
        a, b = a_cores[i], b_cores[i]
        r, s = a.shape[0], b.shape[0]
        c = _product_times_sketch(a, b, w[i + 1])
        cols = c.shape[1]
        bond = min(rank, r * s, cols)
        g = D.to_device(rng.standard_normal((cols, bond)))
        wi = D.empty(r, s, bond)
        D.gemm(1, r * s, bond, cols, 1.0, P(c), cols, 1, 0, P(g), bond, 1, 0, 0.0, P(wi), bond, 1, 0)
        _lib.check(L.ttx_scale_maxabs_f64(D.stream(), r * s * bond, P(wi)), "ttx_scale_maxabs_f64")
        w[i] = wi
    cores = []
    carry = D.to_device(np.ones((1, 1, 1)))
    for i in range(d):
        a, b = a_cores[i], b_cores[i]
        _, m, n, rp = a.shape
        sp = b.shape[3]
        c0 = carry.shape[0]
        mat = _carry_times_product(carry, a, b)          # (c0 m n, rp sp)
        if i == d - 1:
            cores.append(mat.reshape(c0, m, n, rp * sp))
            break
        rows = c0 * m * n
        l1 = w[i + 1].shape[2]
        mw = D.empty(rows, l1)
        D.gemm(1, rows, l1, rp * sp, 1.0, P(mat), rp * sp, 1, 0, P(w[i + 1]), l1, 1, 0, 0.0, P(mw),
               l1, 1, 0)
        k = min(rows, l1)
        Q = D.empty(c0, m, n, k)
        R = D.empty(k, l1)
        D.qr(rows, l1, P(mw), l1, 1, P(Q), k, 1, P(R), l1, 1)
        cores.append(Q)
        nxt = D.empty(k, rp, sp)                         # Q^T mat
        D.gemm(1, k, rp * sp, rows, 1.0, P(Q), 1, k, 0, P(mat), rp * sp, 1, 0, 0.0, P(nxt), rp * sp,
               1, 0)
        carry = nxt
    return cores


def _mirror(c: torch.Tensor) -> torch.Tensor:
    """The core with its bonds swapped, (r, n, m, r') -> (r', n, m, r)."""
    r0, nr, nc, r1 = c.shape
    out = D.empty(r1, nr, nc, r0)
    D.copy_nd([r0, nr, nc, r1], P(c), [nr * nc * r1, nc * r1, r1, 1], P(out),
              [1, nc * r0, r0, nr * nc * r0])
    return out


def randomized_sweep_right(cores: List[torch.Tensor], rank: int, rng: np.random.Generator) -> None:
    """Mirror image of :func:`randomized_sweep_left` (``utils.py:201-227``)."""
    mirrored = [_mirror(c) for c in reversed(cores)]
    randomized_sweep_left(mirrored, rank, rng)
    cores[:] = [_mirror(c) for c in reversed(mirrored)]


# ---------------------------------------------------------------- container
class TensorTrain:
    """Tensor train / tensor train matrix with device cores (``tensor_train.py:150-244``)."""

    def __init__(self, cores: Sequence):
        if len(cores) < 2:
            raise WrongTTFormat("Tensor train must has at least 2 sites")
        shapes = [tuple(c.shape) for c in cores]
        ndims = {len(s) for s in shapes}
        if ndims == {3}:
            tt_type = TensorTrainType.TT
        elif ndims == {4}:
            tt_type = TensorTrainType.TT if all(s[2] == 1 for s in shapes) else TensorTrainType.TTM
        else:
            raise WrongTTFormat("All cores must be 3 or 4 rank tensors")
        bonds, rows, cols, out = [1], [], [], []
        for i, (c, s) in enumerate(zip(cores, shapes)):
            if i == 0 and s[0] != 1:
                raise WrongTTFormat(
                    "Only open boundary conditions is supported, first core must has first index equal to 1")
            if i == len(cores) - 1 and s[-1] != 1:
                raise WrongTTFormat(
                    "Only open boundary conditions is supported, last core must has last index equal to 1")
            if s[0] != bonds[i]:
                raise DimensionMismatch(f"Connected sites bonds are mismatching: {s[0]} != {bonds[i]}")
            t = _as_core(c)
            if len(s) == 3:
                t = t.reshape(s[0], s[1], 1, s[2])
            out.append(t)
            rows.append(s[1])
            cols.append(s[2] if tt_type == TensorTrainType.TTM else 1)
            bonds.append(s[-1])
        self.n_cores = len(out)
        self.cores = out
        self.max_rank = max(bonds)
        self.orth_center = None
        self.row_modes = rows
        self.col_modes = cols
        self.bonds = bonds
        self.tt_type = tt_type

    @classmethod
    def _from_device_cores(cls, cores: Sequence[torch.Tensor], bonds: Sequence[int],
                           modes: Sequence[int]) -> "TensorTrain":
        """A TT over FP64 device cores of shape (bonds[k], modes[k], 1, bonds[k+1])
        that their producer already laid out and checked (the engine's results):
        the constructor's validation and copies are skipped."""
        self = cls.__new__(cls)
        self.n_cores = len(cores)
        self.cores = list(cores)
        self.bonds = list(bonds)
        self.max_rank = max(self.bonds)
        self.orth_center = None
        self.row_modes = list(modes)
        self.col_modes = [1] * len(cores)
        self.tt_type = TensorTrainType.TT
        return self

    # -- bookkeeping
    def __len__(self) -> int:
        return self.n_cores

    def __getitem__(self, idx):
        return self.cores[idx]

    def _sync_bonds(self) -> None:
        self.bonds = [c.shape[0] for c in self.cores] + [1]
        self.max_rank = max(self.bonds)

    def is_ttm(self) -> bool:
        return self.tt_type == TensorTrainType.TTM

    def physical_dims(self):
        if self.tt_type is TensorTrainType.TT:
            return list(self.row_modes)
        return list(zip(self.row_modes, self.col_modes))

    def copy(self) -> "TensorTrain":
        t = TensorTrain.__new__(TensorTrain)
        t.n_cores = self.n_cores
        t.cores = [c.clone() for c in self.cores]
        t.max_rank = self.max_rank
        t.orth_center = self.orth_center
        t.row_modes = list(self.row_modes)
        t.col_modes = list(self.col_modes)
        t.bonds = list(self.bonds)
        t.tt_type = self.tt_type
        return t

    def to_numpy(self) -> List[np.ndarray]:
        """Host copies of the cores."""
        return [c.detach().cpu().numpy() for c in self.cores]

    # -- canonical forms
    def orthonormalize(self, canonical: str = "left", center_position: Optional[int] = None,
                       inplace: bool = False) -> "TensorTrain":
        """QR sweep with the reference's centre bookkeeping (``tensor_train.py:705-787``)."""
        d = self.n_cores
        if center_position is not None and not 0 <= center_position < d:
            raise ValueError(
                f"center_position must be in [0, {d - 1}], but got {center_position}")
        t = self if inplace else self.copy()
        had = t.orth_center is not None
        if canonical == "left":
            end = center_position if center_position is not None else d - 1
            start = 0
            if t.orth_center is not None:
                start = t.orth_center
                t.orth_center = end if start <= end else t.orth_center
            if start >= end:
                return t
            qr_sweep_left(t.cores, start, end)
        elif canonical == "right":
            end = center_position if center_position is not None else 0
            start = d - 1
            if t.orth_center is not None:
                start = t.orth_center
                t.orth_center = end if start >= end else t.orth_center
            if start <= end:
                return t
            qr_sweep_right(t.cores, start, end)
        elif canonical == "mix":
            if center_position is None:
                center_position = d // 2
            end = center_position
            qr_sweep_left(t.cores, 0, center_position)
            qr_sweep_right(t.cores, d - 1, center_position)
        else:
            raise ValueError(f"Canonical form must be 'left', 'right' or 'mix', but got {canonical}")
        t.orth_center = end if (had or canonical == "mix" or end in (0, d - 1)) else None
        t._sync_bonds()
        return t

    def compress_svd(self, error: float = 0.0, max_rank: Optional[int] = None, canonical: str = "auto",
                     inplace: bool = False, spectra=None) -> "TensorTrain":
        """SVD sweep, direction from the centre (``tensor_train.py:789-850``)."""
        t = self if inplace else self.copy()
        if canonical == "auto":
            canonical = "left" if (t.orth_center is None or t.orth_center <= (t.n_cores - 1) / 2) else "right"
        if canonical == "right":
            svd_sweep_right(t.cores, error, max_rank, spectra)
            t.orth_center = 0
        elif canonical == "left":
            svd_sweep_left(t.cores, error, max_rank, spectra)
            t.orth_center = t.n_cores - 1
        else:
            raise ValueError(f"Canonical form must be 'left', 'right' or 'auto', but got {canonical}")
        t._sync_bonds()
        return t

    def truncate(self, error: float = 0.0, max_rank: Optional[int] = None, canonical: str = "auto",
                 inplace: bool = False, method: str = "direct", seed: Optional[int] = None,
                 options: Optional[dict] = None, spectra=None,
                 precision: str = "fp64") -> "TensorTrain":
        """Round to minimal ranks: QR sweep then SVD sweep (``tensor_train.py:852-940``).

        ``method="randomized"`` replaces the QR sweep by the randomized sweep
        (:func:`randomized_sweep_left` or its mirror image) on a sketch of bond
        ``max_rank + options["oversampling"]`` (default 10) drawn from
        ``default_rng(seed)``; it needs ``max_rank`` and ``error`` bounds the
        SVD truncation only.  ``precision="mixed"`` runs the absorb
        contractions of both sweeps in FP32 (FP32 MFMA, FP32 accumulation) and
        keeps the QR and SVD in FP64.
        """
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        if canonical not in ("auto", "left", "right"):
            raise ValueError(f"Canonical form must be 'left', 'right' or 'auto', but got {canonical}")
        if method not in ("direct", "randomized"):
            raise ValueError(f"Method must be 'direct' or 'randomized', but got {method!r}")
        if method == "randomized":
            rank = _sketch_rank(max_rank, options)
        if max_rank is not None and max_rank <= 0:
            raise ValueError("max_rank must be a positive integer.")
        t = self if inplace else self.copy()
        form = "left" if canonical == "right" else "right"
        if canonical == "auto":
            form = "left" if (t.orth_center is None or t.orth_center >= t.n_cores / 2) else "right"
        with D.contractions("fp32" if precision == "mixed" else "fp64"):
            if method == "direct":
                t.orthonormalize(canonical=form, inplace=True)
            else:
                rng = np.random.default_rng(seed)
                if form == "left":
                    randomized_sweep_left(t.cores, rank, rng)
                    t.orth_center = t.n_cores - 1
                else:
                    randomized_sweep_right(t.cores, rank, rng)
                    t.orth_center = 0
                t._sync_bonds()
            t.compress_svd(error, max_rank, inplace=True, spectra=spectra)
        return t

    # -- arithmetic used by the residual check
    def add(self, other: "TensorTrain", truncate: bool = False, error=None, max_rank=None,
            method: str = "direct", seed=None, options=None) -> "TensorTrain":
        """Block-placed sum (``tensor_train.py:254-322``)."""
        if not isinstance(other, TensorTrain):
            raise TypeError(f"unsupported operand type(s) for +: 'TensorTrain' and '{type(other)}'")
        if self.physical_dims() != other.physical_dims():
            raise DimensionMismatch("Tensor trains must have the same dimensions")
        d = self.n_cores
        cores = []
        for i in range(d):
            a, b = self.cores[i], other.cores[i]
            L = 1 if i == 0 else a.shape[0] + b.shape[0]
            R = 1 if i == d - 1 else a.shape[3] + b.shape[3]
            nr, nc = a.shape[1], a.shape[2]
            c = D.zeros(L, nr, nc, R)
            st = [nr * nc * R, nc * R, R, 1]
            D.copy_nd(list(a.shape), P(a), [nr * nc * a.shape[3], nc * a.shape[3], a.shape[3], 1], P(c), st)
            off = (L - b.shape[0]) * nr * nc * R + (R - b.shape[3])
            D.copy_nd(list(b.shape), P(b), [nr * nc * b.shape[3], nc * b.shape[3], b.shape[3], 1], P(c, off), st)
            cores.append(c)
        res = TensorTrain(cores)
        return _round(res, truncate, error, max_rank, method, seed, options)

    def __add__(self, other):
        return self.add(other)

    def multiply(self, other, truncate: bool = False, error=None, max_rank=None, method="direct",
                 seed=None, options=None) -> "TensorTrain":
        """Element-wise product with a train (cores of bonds r s, ``tensor_train.py:363-395``)
        or a scalar multiple on the centre core, or core 0 (``tensor_train.py:396-401``)."""
        if isinstance(other, TensorTrain):
            if self.physical_dims() != other.physical_dims():
                raise DimensionMismatch("Tensor trains must have the same dimensions")
            if truncate and method == "randomized":
                prod = TensorTrain(randomized_hadamard(self.cores, other.cores,
                                                       _sketch_rank(max_rank, options),
                                                       np.random.default_rng(seed)))
                prod.orth_center = prod.n_cores - 1
                return prod.compress_svd(0.0 if error is None else error, max_rank, inplace=True)
            cores = []
            for a, b in zip(self.cores, other.cores):
                ra, ni, nj, rb = a.shape
                sa, sb = b.shape[0], b.shape[3]
                c = D.empty(ra * sa, ni, nj, rb * sb)
                _lib.check(D.lib().ttx_hadamard_core_f64(D.stream(), P(a), ra, ni, nj, rb, P(b), sa,
                                                         sb, P(c)), "ttx_hadamard_core_f64")
                cores.append(c)
            return _round(TensorTrain(cores), truncate, error, max_rank, method, seed, options)
        if not isinstance(other, numbers.Number):
            raise TypeError(f"unsupported operand type(s) for *: 'TensorTrain' and '{type(other)}'")
        res = self.copy()
        k = res.orth_center if res.orth_center is not None else 0
        c = res.cores[k]
        D.copy_nd([c.numel()], P(c), [1], P(c), [1], float(other))
        return _round(res, truncate, error, max_rank, method, seed, options)

    def __mul__(self, other):
        return self.multiply(other)

    def __rmul__(self, other):
        return self.multiply(other)

    def __neg__(self):
        return self.multiply(-1)

    def __sub__(self, other):
        return self.add(-other)

    def dot(self, other: "TensorTrain", truncate: bool = False, error=None, max_rank=None,
            method="direct", seed=None, options=None) -> "TensorTrain":
        """Matrix product of two trains, cores ``(Ra r, n_row, m_col, Ra' r')``
        (``tensor_train.py:427-499``): TTM @ TT on the matvec kernel, TTM @ TTM,
        TT @ TT^T and TT^T @ TT on the general core-product kernel."""
        if not isinstance(other, TensorTrain):
            raise TypeError(f"unsupported operand type(s) for @: 'TensorTrain' and '{type(other)}'")
        if self.col_modes != other.row_modes:
            raise DimensionMismatch(
                "Tensor trains must have the same last and first physical dimensions respectively")
        cores = []
        for a, x in zip(self.cores, other.cores):
            Ra, nr, nk, Rb = a.shape
            rx, _, nc, rx2 = x.shape
            c = D.empty(Ra * rx, nr, nc, Rb * rx2)
            if nc == 1:
                D.ttm_core(a, x, P(c), nr * Rb * rx2, Rb * rx2, 1)
            else:
                _lib.check(D.lib().ttx_ttm_ttm_core_f64(D.stream(), P(a), Ra, nr, nk, Rb, P(x), rx,
                                                        nc, rx2, P(c)), "ttx_ttm_ttm_core_f64")
            cores.append(c)
        return _round(TensorTrain(cores), truncate, error, max_rank, method, seed, options)

    def transpose(self, cores_ind: Optional[List[int]] = None, conjugate: bool = False,
                  inplace: bool = False) -> "TensorTrain":
        """Swap row and column modes of the given cores, all by default
        (``tensor_train.py:647-689``); cores are real, so ``conjugate`` changes nothing."""
        t = self if inplace else self.copy()
        idx = range(t.n_cores) if cores_ind is None else cores_ind
        for i in idx:
            c = t.cores[i]
            r, nr, nc, r1 = c.shape
            out = D.empty(r, nc, nr, r1)
            D.copy_nd([r, nc, nr, r1], P(c), [nr * nc * r1, r1, nc * r1, 1], P(out),
                      [nc * nr * r1, nr * r1, r1, 1])
            t.cores[i] = out
            t.row_modes[i], t.col_modes[i] = nc, nr
        ttm = any(m > 1 for m in t.col_modes)
        t.tt_type = TensorTrainType.TTM if ttm else TensorTrainType.TT
        return t

    def __matmul__(self, other):
        return self.dot(other)

    def tt_norm_device(self) -> torch.Tensor:
        """Frobenius norm as a device scalar (``tensor_train.py:509-540``)."""
        if self.orth_center is not None:
            return D.norm(self.cores[self.orth_center])
        cores = [c.reshape(c.shape[0], c.shape[1] * c.shape[2], 1, c.shape[3]) for c in self.cores]
        cores = list(cores)
        qr_sweep_right(cores, len(cores) - 1, 0, form_q=False)
        return D.norm(cores[0])

    def tt_norm(self) -> float:
        return float(self.tt_norm_device().item())

    def get_full_tensor(self) -> np.ndarray:
        """Dense tensor of a small train, contracted on the device: shape (rows...)
        for a TT, (rows..., cols...) for a TTM (``tensor_train.py:599-624``)."""
        t = self.cores[0].reshape(-1, self.cores[0].shape[3])
        for c in self.cores[1:]:
            m, r = t.shape
            rest = c.numel() // c.shape[0]
            out = D.empty(m, rest)
            D.gemm(1, m, rest, r, 1.0, P(t), r, 1, 0, P(c), rest, 1, 0, 0.0, P(out), rest, 1, 0)
            t = out.reshape(-1, c.shape[3])
        if not self.is_ttm():
            return t.reshape(self.row_modes).cpu().numpy()
        d = self.n_cores
        shape = [x for pair in zip(self.row_modes, self.col_modes) for x in pair]
        perm = list(range(0, 2 * d, 2)) + list(range(1, 2 * d, 2))
        return np.ascontiguousarray(t.reshape(shape).cpu().numpy().transpose(perm))

    def matricize(self) -> np.ndarray:
        """The full tensor as a vector (TT) or a (prod rows, prod cols) matrix (TTM)
        (``tensor_train.py:626-638``)."""
        m = _merge_group(self.cores)
        if not self.is_ttm():
            return m.reshape(-1).cpu().numpy()
        return m.reshape(m.shape[1], m.shape[2]).cpu().numpy()

    # -- QTT conversions
    def tt_to_qtt(self, mode_size: int = 2, error: float = 1e-12,
                  max_rank: Optional[int] = None) -> "TensorTrain":
        """Quantized train (``tensor_train.py:942-1073``): a core of mode
        ``mode_size**q`` becomes ``q`` cores of mode ``mode_size``.

        After a right QR sweep, each split is a truncated SVD of the unfolding
        (left, m, m_c) x (rest) with delta = error / sqrt(#SVDs) * ||x||; the kept
        U is emitted and S V^T carried on (the reference carries U^T M, equal to
        it up to rounding); at a core boundary one more SVD pushes the remainder
        into the next core.  Left-canonical, centre at the last core; the
        receiver is unchanged."""
        if mode_size < 2:
            raise ValueError(f"mode_size must be at least 2, but got {mode_size}")
        if max_rank is not None and max_rank <= 0:
            raise ValueError("max_rank must be a positive integer.")
        qs = []
        for i, row in enumerate(self.row_modes):
            q = 0
            while mode_size ** q < row:
                q += 1
            if mode_size ** q != row:
                raise DimensionMismatch(f"Row mode {row} not a power of {mode_size}")
            if self.is_ttm() and mode_size ** q != self.col_modes[i]:
                raise DimensionMismatch("Column mode size of TTM not a power of mode size")
            qs.append(q)
        rs, cs = mode_size, (mode_size if self.is_ttm() else 1)
        n_svd = max(1, sum(max(0, q - 1) for q in qs) + self.n_cores - 1)
        tt = self.orthonormalize(canonical="right")
        nrm = tt.tt_norm_device()
        scale = error / math.sqrt(n_svd)
        buf = torch.zeros(2, dtype=torch.int32, device=D.device())
        out = []
        for i, q in enumerate(qs):
            cur = tt.cores[i]
            left, right = cur.shape[0], cur.shape[3]
            if q == 0:
                raise ValueError(f"cannot split mode 1 of core {i} into modes of {mode_size}")
            for j in range(q - 1, 0, -1):
                R, C = rs ** j, cs ** j                   # what stays behind this split
                rows, cols = left * rs * cs, R * C * right
                mat = cur
                if cs > 1:                                # (left, rs, R, cs, C, right) -> (left, rs, cs, R, C, right)
                    mat = D.empty(rows, cols)
                    D.copy_nd([left * rs, cs, R, C * right], P(cur),
                              [R * cs * C * right, C * right, cs * C * right, 1], P(mat),
                              [cs * R * C * right, R * C * right, C * right, 1])
                u, sv, nb = _svd_split(rows, cols, mat, nrm, scale, max_rank, buf, i)
                out.append(u.reshape(left, rs, cs, nb))
                cur = sv.reshape(nb, R, C, right)
                left = nb
            if i < self.n_cores - 1:
                rows = left * rs * cs
                u, sv, nb = _svd_split(rows, right, cur, nrm, scale, max_rank, buf, i)
                out.append(u.reshape(left, rs, cs, nb))
                tt.cores[i + 1] = D.absorb_left(P(sv), nb, right, tt.cores[i + 1])
            else:
                out.append(cur.reshape(left, rs, cs, right))
        qtt = TensorTrain(out)
        qtt.orth_center = qtt.n_cores - 1
        return qtt

    def qtt_to_tt(self, physical_indexes, error: float = 1e-12,
                  max_rank: Optional[int] = None) -> "TensorTrain":
        """Merge runs of consecutive cores into cores of the given modes, then
        ``truncate(error, max_rank)`` (``tensor_train.py:1075-1135``)."""
        counts, pos = [], 0
        for i, idx in enumerate(physical_indexes):
            row_t, col_t = idx if self.is_ttm() else (idx, 1)
            pr = pc = 1
            n = 0
            while n == 0 or pr != row_t or pc != col_t:
                if pos + n >= self.n_cores:
                    raise DimensionMismatch(
                        f"Not enough QTT cores to match physical index for core {i}")
                pr *= self.row_modes[pos + n]
                pc *= self.col_modes[pos + n]
                n += 1
                if pr > row_t:
                    raise DimensionMismatch(
                        f"Row mode size of QTT exceeds original row mode size for core {i}")
                if pc > col_t:
                    raise DimensionMismatch(
                        f"Column mode size of QTT exceeds original column mode size for core {i}")
            counts.append(n)
            pos += n
        if pos != self.n_cores:
            raise DimensionMismatch(f"Physical indexes cover only {pos} of {self.n_cores} QTT cores")
        cores, pos = [], 0
        for n in counts:
            cores.append(_merge_group(self.cores[pos:pos + n]))
            pos += n
        return TensorTrain(cores).truncate(error=error, max_rank=max_rank, inplace=True)


def _merge_group(cores: Sequence[torch.Tensor]) -> torch.Tensor:
    """Consecutive cores as one core (left, prod rows, prod cols, right), rows and
    columns each in the cores' order (``tensor_train.py:58-91``): one GEMM per
    core, and where both the accumulated column mode and the new row mode exceed
    1, a strided copy that moves the new row index ahead of the column index."""
    m = cores[0]
    if len(cores) == 1:
        return m.clone()
    for c in cores[1:]:
        left, R, C, mid = m.shape
        _, r, cc, right = c.shape
        rest = r * cc * right
        prod = D.empty(left, R, C, r, cc, right)
        D.gemm(1, left * R * C, rest, mid, 1.0, P(m), mid, 1, 0, P(c), rest, 1, 0, 0.0, P(prod),
               rest, 1, 0)
        if C > 1 and r > 1:
            swapped = D.empty(left, R, r, C, cc, right)
            D.copy_nd([left * R, r, C, cc * right], P(prod), [C * rest, cc * right, rest, 1],
                      P(swapped), [r * C * cc * right, C * cc * right, cc * right, 1])
            prod = swapped
        m = prod.reshape(left, R * r, C * cc, right)
    return m


def _svd_split(rows: int, cols: int, mat: torch.Tensor, nrm: torch.Tensor, scale: float,
               max_rank, buf: torch.Tensor, core: int):
    """Truncated SVD of a row-major (rows x cols) matrix with delta = scale * nrm:
    returns the kept U (rows x k), S V^T (k x cols) and k."""
    U, S, Vt = D.svd(rows, cols, P(mat), cols, 1, buf, 0)
    k = min(rows, cols)
    D.truncated_bond_dev(S, nrm, scale, max_rank, buf, 1)
    nb = _sweep_status(buf, core)
    u = D.empty(rows, nb)
    D.copy_nd([rows, nb], P(U), [k, 1], P(u), [nb, 1])
    D.scale_cols(cols, nb, P(Vt), 1, cols, P(S))
    return u, Vt[:nb], nb


def _round(res: TensorTrain, truncate: bool, error, max_rank, method, seed=None,
           options=None) -> TensorTrain:
    """``_round_result`` (``tensor_train.py:94-119``)."""
    if not truncate:
        if method != "direct" or error is not None or max_rank is not None:
            raise ValueError("error, max_rank and method need truncate=True")
        return res
    return res.truncate(error=0.0 if error is None else error, max_rank=max_rank, inplace=True,
                        method=method, seed=seed, options=options)


def _ranks(d: int, ranks) -> List[int]:
    if not isinstance(ranks, list):
        return [1] + [int(ranks)] * (d - 1) + [1]
    return list(ranks)


def ones(physical_indexes, ranks=1) -> TensorTrain:
    """All-ones train (``tensor_train.py:1250-1285``)."""
    if len(physical_indexes) < 2:
        raise WrongTTFormat("Tensor train must have at least two cores.")
    r = _ranks(len(physical_indexes), ranks)
    cores = []
    for i, p in enumerate(physical_indexes):
        nr, nc = (p, 1) if isinstance(p, (int, np.integer)) else (p[0], p[1])
        c = D.empty(r[i], nr, nc, r[i + 1])
        D.fill(c, 1.0)
        cores.append(c)
    return TensorTrain(cores)


def zeros(physical_indexes, ranks=1) -> TensorTrain:
    """All-zeros train (``tensor_train.py:1288-1318``)."""
    t = ones(physical_indexes, ranks)
    for c in t.cores:
        D.fill(c, 0.0)
    return t


def eye(physical_indexes) -> TensorTrain:
    """Identity TTM of bond 1 (``tensor_train.py:1321-1335``)."""
    if len(physical_indexes) < 2:
        raise WrongTTFormat("Tensor train must have at least two cores.")
    return TensorTrain([np.eye(p).reshape(1, p, p, 1) for p in physical_indexes])


def rand(physical_indexes, ranks=1, seed=None) -> TensorTrain:
    """Uniform [0, 1) cores in the reference's draw order (``tensor_train.py:1350-1387``).

    The numbers come from NumPy's PCG64 ``default_rng(seed)`` on the host (the
    same generator the reference calls), then live on the device.
    """
    if len(physical_indexes) < 2:
        raise WrongTTFormat("Tensor train must have at least two cores.")
    rng = np.random.default_rng(seed)
    r = _ranks(len(physical_indexes), ranks)
    cores = []
    for i, p in enumerate(physical_indexes):
        shape = (r[i], p, 1, r[i + 1]) if isinstance(p, (int, np.integer)) else (r[i], p[0], p[1], r[i + 1])
        cores.append(rng.random(shape))
    return TensorTrain(cores)
