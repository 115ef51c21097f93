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
