"""Input builders for the solver workloads: QTT operators and right-hand sides.

These produce plain host arrays (the core list of a train, each core indexed
``(left_bond, row_mode, col_mode, right_bond)``).  They are problem
construction, which the reference keeps outside the timed region
(``benchmarks/poisson.py:192-203``); the solve itself runs on the device.

Operators:

* :func:`laplacian_1d_cores` -- Dirichlet ``-d^2/dx^2`` on ``2**n_bits``
  interior points in QTT form, most significant bit first, TTM bond 3.  The
  construction follows the shift-operator form the reference's own example
  writes down (``examples/poisson_2d.py:36-50``): a bit pattern of the index
  with a carry chain for ``i -> i+1`` and ``i -> i-1``.
* :func:`poisson_2d_cores` -- the Kronecker sum ``L (x) I + I (x) L`` with the
  x bits first, formed as the block-placed sum of two trains exactly as
  ``TensorTrain.add`` does (``tensor_train.py:254-322``), giving TTM bonds
  ``[1, 4, ..., 4, 2, 4, ..., 4, 1]`` (``examples/poisson_2d.py:58-60``).
* :func:`shifted_laplacian_sum_cores` -- ``sum_i (L_i + c I)`` on ``d`` modes of
  size ``n`` as an exact rank-2 TTM.
* :func:`diffusion_2d_cores` -- ``-div(k grad u)`` on a ``2**L`` square grid with
  ``k(x, y) = 1 + theta sin(pi x) sin(pi y)`` (five-point finite volumes,
  Dirichlet, ``h = 1/(2**L + 1)``), x bits first.  With ``s = sin(pi .)`` the
  coefficient is separable, so ``A(theta) = A0 + theta A1`` with
  ``A0 = L (x) I + I (x) L`` and ``A1 = M (x) S + S (x) M`` where ``S`` is
  ``diag(s)`` at the nodes and ``M = D^T diag(s) D / h^2`` the 1-D stiffness
  with ``s`` at the cell edges.  ``S`` and ``M`` are put in QTT form by a
  TT-SVD of their dense ``2**L x 2**L`` matrices; the sum is rounded.
"""

from __future__ import annotations

from typing import List, Sequence

import numpy as np

_I2 = np.eye(2)
# one bit of i -> i + 1 without carry: |1><0| acting on the column index
_J = np.array([[0.0, 1.0], [0.0, 0.0]])
_O = np.zeros((2, 2))


def _core_from_blocks(blocks: Sequence[Sequence[np.ndarray]]) -> np.ndarray:
    rows = len(blocks)
    cols = len(blocks[0])
    n = blocks[0][0].shape[0]
    core = np.zeros((rows, n, n, cols))
    for a in range(rows):
        for c in range(cols):
            core[a, :, :, c] = blocks[a][c]
    return core


def laplacian_1d_cores(n_bits: int, h: float) -> List[np.ndarray]:
    """QTT cores of ``tridiag(-1, 2, -1) / h^2`` of size ``2**n_bits``.

    Bond index 0 = no carry, 1 = the up-shift carries, 2 = the down-shift
    carries (the carry travels from the least significant bit, which is the
    last core, towards the first).
    """
    if n_bits < 2:
        raise ValueError("n_bits must be at least 2")
    first = _core_from_blocks([[_I2, _J, _J.T]])
    middle = _core_from_blocks([[_I2, _J, _J.T], [_O, _J.T, _O], [_O, _O, _J]])
    last = _core_from_blocks([[2 * _I2 - _J - _J.T], [-_J.T], [-_J]]) / h**2
    return [first] + [middle.copy() for _ in range(n_bits - 2)] + [last]


def identity_cores(n_bits: int, n: int = 2) -> List[np.ndarray]:
    return [np.eye(n).reshape(1, n, n, 1) for _ in range(n_bits)]


def block_sum_cores(a: List[np.ndarray], b: List[np.ndarray]) -> List[np.ndarray]:
    """Cores of the sum of two trains by block placement (``tensor_train.py:254-322``).

    Interior bonds add; the first core concatenates along the right bond and
    the last along the left bond.
    """
    d = len(a)
    out = []
    for i, (ca, cb) in enumerate(zip(a, b)):
        la, ra = ca.shape[0], ca.shape[3]
        lb, rb = cb.shape[0], cb.shape[3]
        left = 1 if i == 0 else la + lb
        right = 1 if i == d - 1 else ra + rb
        core = np.zeros((left, ca.shape[1], ca.shape[2], right), dtype=np.result_type(ca, cb))
        core[:la, :, :, :ra] = ca
        core[left - lb:, :, :, right - rb:] = cb
        out.append(core)
    return out


def poisson_2d_cores(n_bits: int) -> List[np.ndarray]:
    """``-Laplace`` on a ``2**n_bits`` square grid, Dirichlet, ``h = 1/(2**n_bits + 1)``.

    x bits first, then y bits; TTM bonds ``[1, 4, ..., 4, 2, 4, ..., 4, 1]``.
    """
    h = 1.0 / (2**n_bits + 1)
    lap = laplacian_1d_cores(n_bits, h)
    eye = identity_cores(n_bits)
    return block_sum_cores(lap + eye, eye + lap)


def shifted_laplacian_sum_cores(d: int, n: int, shift: float) -> List[np.ndarray]:
    """``sum_i (L_i + shift * I)`` with ``L = tridiag(-1, 2, -1)`` of size ``n``.

    Exact rank-2 TTM: first ``[M, I]``, middle ``[[I, 0], [M, I]]``, last
    ``[I; M]`` with ``M = L + shift * I``.
    """
    L = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    M = L + shift * np.eye(n)
    I = np.eye(n)
    Z = np.zeros((n, n))
    first = _core_from_blocks([[M, I]])
    middle = _core_from_blocks([[I, Z], [M, I]])
    last = _core_from_blocks([[I], [M]])
    return [first] + [middle.copy() for _ in range(d - 2)] + [last]


def ttm_cores_from_dense(mat: np.ndarray, n_bits: int, tol: float = 1e-14) -> List[np.ndarray]:
    """QTT (mode 2 x 2) cores of a dense ``2**n_bits x 2**n_bits`` matrix by TT-SVD.

    Row and column indices are split into bits, most significant first; core
    ``k`` carries bit ``k`` of the row and of the column.  Singular values
    below ``tol`` times the largest of each unfolding are dropped.
    """
    n = 2**n_bits
    if mat.shape != (n, n):
        raise ValueError(f"expected a {n} x {n} matrix")
    t = mat.reshape([2] * (2 * n_bits))
    order = [x for k in range(n_bits) for x in (k, n_bits + k)]
    rest = np.ascontiguousarray(t.transpose(order)).reshape(1, -1)
    cores, r = [], 1
    for _ in range(n_bits - 1):
        u, sv, vt = np.linalg.svd(rest.reshape(r * 4, -1), full_matrices=False)
        keep = max(1, int(np.count_nonzero(sv > tol * sv[0]))) if sv[0] > 0 else 1
        cores.append(u[:, :keep].reshape(r, 2, 2, keep))
        rest = sv[:keep, None] * vt[:keep]
        r = keep
    cores.append(rest.reshape(r, 2, 2, 1))
    return cores


def round_cores(cores: List[np.ndarray], tol: float = 1e-14) -> List[np.ndarray]:
    """TT rounding of a train's cores (host, builder use): right-to-left QR,
    then left-to-right SVD keeping singular values above ``tol * sigma_1``."""
    cores = [c.copy() for c in cores]
    d = len(cores)
    for k in range(d - 1, 0, -1):
        r, a, b, r2 = cores[k].shape
        q, rr = np.linalg.qr(cores[k].reshape(r, -1).T)
        cores[k] = q.T.reshape(-1, a, b, r2)
        cores[k - 1] = np.tensordot(cores[k - 1], rr.T, axes=(3, 0))
    for k in range(d - 1):
        r, a, b, r2 = cores[k].shape
        u, sv, vt = np.linalg.svd(cores[k].reshape(-1, r2), full_matrices=False)
        keep = max(1, int(np.count_nonzero(sv > tol * sv[0]))) if sv[0] > 0 else 1
        cores[k] = u[:, :keep].reshape(r, a, b, keep)
        carry = sv[:keep, None] * vt[:keep]
        cores[k + 1] = np.tensordot(carry, cores[k + 1], axes=(1, 0))
    return cores


def _scaled_sum(parts):
    """Cores of sum_i c_i * T_i for [(c_i, cores_i)] (block placement)."""
    out = None
    for c, cores in parts:
        cores = [x.copy() for x in cores]
        cores[0] = cores[0] * c
        out = cores if out is None else block_sum_cores(out, cores)
    return out


def diffusion_2d_parts(n_bits: int, tol: float = 1e-14):
    """``(A0, A1)`` with ``A(theta) = A0 + theta A1`` (see :func:`diffusion_2d_cores`)."""
    n = 2**n_bits
    h = 1.0 / (n + 1)
    x_nodes = (np.arange(n) + 1) * h
    x_edges = (np.arange(n + 1) + 0.5) * h
    s_nodes = np.sin(np.pi * x_nodes)
    s_edges = np.sin(np.pi * x_edges)
    # M = D^T diag(s_e) D / h^2 with D the (n+1) x n difference matrix: tridiagonal
    main = (s_edges[:-1] + s_edges[1:]) / h**2
    off = -s_edges[1:-1] / h**2
    M = np.diag(main) + np.diag(off, 1) + np.diag(off, -1)
    lap = laplacian_1d_cores(n_bits, h)
    eye = identity_cores(n_bits)
    S = ttm_cores_from_dense(np.diag(s_nodes), n_bits, tol)
    Mq = ttm_cores_from_dense(M, n_bits, tol)
    A0 = round_cores(block_sum_cores(lap + eye, eye + lap), tol)
    A1 = round_cores(block_sum_cores(Mq + S, S + Mq), tol)
    return A0, A1


def diffusion_2d_cores(n_bits: int, theta: float, parts=None, tol: float = 1e-13) -> List[np.ndarray]:
    """``-div(k grad u)`` on a ``2**n_bits`` square grid, ``k = 1 + theta sin(pi x) sin(pi y)``.

    ``parts`` may pass a cached :func:`diffusion_2d_parts` result for the same
    ``n_bits``.
    """
    A0, A1 = parts if parts is not None else diffusion_2d_parts(n_bits)
    return round_cores(_scaled_sum([(1.0, A0), (float(theta), A1)]), tol)


def random_train_cores(modes: Sequence[int], ranks, seed, normal: bool = False) -> List[np.ndarray]:
    """Cores drawn in the order and shapes of the reference's ``rand``.

    ``tensor_train.py:1350-1387``: one ``default_rng(seed)`` generator, cores
    drawn site by site with shape ``(r_i, n_i, 1, r_{i+1})``; ``random`` gives
    uniform ``[0, 1)`` (the reference's ``rand``), ``normal=True`` draws
    ``standard_normal`` in the same order.
    """
    rng = np.random.default_rng(seed)
    d = len(modes)
    if not isinstance(ranks, (list, tuple)):
        ranks = [1] + [int(ranks)] * (d - 1) + [1]
    draw = rng.standard_normal if normal else rng.random
    return [draw((ranks[i], modes[i], 1, ranks[i + 1])) for i in range(d)]


def ones_cores(modes: Sequence[int], ranks=1) -> List[np.ndarray]:
    """Cores of ``ones`` (``tensor_train.py:1250-1285``)."""
    d = len(modes)
    if not isinstance(ranks, (list, tuple)):
        ranks = [1] + [int(ranks)] * (d - 1) + [1]
    return [np.ones((ranks[i], modes[i], 1, ranks[i + 1])) for i in range(d)]
