"""Exact discrete solution of the 2-D Dirichlet Poisson problem of C1.

The QTT operator of ``ttnumpy_amd.problems.poisson_2d_cores(n_bits)`` (the
form of ``examples/poisson_2d.py:36-60``) is ``A = (T (x) I + I (x) T) / h^2``
with ``T = tridiag(-1, 2, -1)`` of size ``n = 2**n_bits``, ``h = 1/(n+1)``,
x bits (most significant first) before y bits, so a full tensor reshaped to
``(n, n)`` is ``U[ix, iy]``.  The type-I discrete sine transform ``S``
(orthonormal, ``S = S^T = S^-1``) diagonalizes ``T``:
``T = S diag(2 - 2 cos(k pi/(n+1))) S``, so ``A^-1 b = S (S B S ./ Lam) S``.
A is SPD with smallest eigenvalue ``lam_min = 2 (2 - 2 cos(pi/(n+1))) / h^2``,
so any x satisfies ``||x - A^-1 b|| <= ||A x - b|| / lam_min``: a bound on a
solver's solution error that involves no other solver.
"""

from __future__ import annotations

import numpy as np
from scipy.fft import dstn, idstn


def eigenvalues_1d(n: int) -> np.ndarray:
    k = np.arange(1, n + 1)
    return 2.0 - 2.0 * np.cos(k * np.pi / (n + 1))


def lam_min(n_bits: int) -> float:
    n = 2**n_bits
    h = 1.0 / (n + 1)
    return 2.0 * eigenvalues_1d(n)[0] / h**2


def kappa(n_bits: int) -> float:
    ev = eigenvalues_1d(2**n_bits)
    return float(ev[-1] / ev[0])


def solve(B: np.ndarray) -> np.ndarray:
    """``A^-1 b`` for a right-hand side given as the ``(n, n)`` array ``B[ix, iy]``."""
    n = B.shape[0]
    h = 1.0 / (n + 1)
    ev = eigenvalues_1d(n) / h**2
    lam = ev[:, None] + ev[None, :]
    return idstn(dstn(B, type=1, norm="ortho") / lam, type=1, norm="ortho")


def apply(U: np.ndarray) -> np.ndarray:
    """``A u`` by the 5-point stencil, for ``U[ix, iy]`` of shape ``(n, n)``."""
    n = U.shape[0]
    h = 1.0 / (n + 1)
    P = np.zeros((n + 2, n + 2))
    P[1:-1, 1:-1] = U
    return (4.0 * U - P[:-2, 1:-1] - P[2:, 1:-1] - P[1:-1, :-2] - P[1:-1, 2:]) / h**2


def error_bound(x_full: np.ndarray, b_full: np.ndarray, n_bits: int):
    """``(||x - A^-1 b||, ||A x - b|| / lam_min, ||A^-1 b||)`` for full tensors."""
    n = 2**n_bits
    X = np.asarray(x_full).reshape(n, n)
    B = np.asarray(b_full).reshape(n, n)
    exact = solve(B)
    err = float(np.linalg.norm(X - exact))
    bound = float(np.linalg.norm(apply(X) - B)) / lam_min(n_bits)
    return err, bound, float(np.linalg.norm(exact))


def diffusion_sparse(n_bits: int, theta: float):
    """The operator of ``problems.diffusion_2d_cores(n_bits, theta)`` (C4's
    ``-div(k grad u)``, ``k = 1 + theta sin(pi x) sin(pi y)``) as a sparse
    matrix in the same index order: ``(T (x) I + I (x) T) / h^2 + theta (M (x) S
    + S (x) M)`` with ``M = D^T diag(sin(pi x_edges)) D / h^2`` and
    ``S = diag(sin(pi x_nodes))``.  ``M (x) S + S (x) M`` is positive
    semidefinite for ``theta >= 0``, so ``lam_min`` of the Laplacian alone
    bounds this operator's smallest eigenvalue from below."""
    import scipy.sparse as sp
    n = 2**n_bits
    h = 1.0 / (n + 1)
    x_nodes = (np.arange(n) + 1) * h
    x_edges = (np.arange(n + 1) + 0.5) * h
    s_nodes = np.sin(np.pi * x_nodes)
    s_edges = np.sin(np.pi * x_edges)
    T = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]) / h**2
    M = sp.diags([-s_edges[1:-1], s_edges[:-1] + s_edges[1:], -s_edges[1:-1]], [-1, 0, 1]) / h**2
    S = sp.diags(s_nodes)
    I = sp.identity(n)
    return (sp.kron(T, I) + sp.kron(I, T) + theta * (sp.kron(M, S) + sp.kron(S, M))).tocsc()


def diffusion_error_bound(x_full: np.ndarray, n_bits: int, theta: float):
    """``(||x - A^-1 b||, ||A x - b|| / lam_min(-Laplace), ||A^-1 b||)`` for
    C4's operator and ``b = ones``."""
    import scipy.sparse.linalg as spla
    A = diffusion_sparse(n_bits, theta)
    b = np.ones(A.shape[0])
    exact = spla.spsolve(A, b)
    x = np.asarray(x_full).reshape(-1)
    err = float(np.linalg.norm(x - exact))
    bound = float(np.linalg.norm(A @ x - b)) / lam_min(n_bits)
    return err, bound, float(np.linalg.norm(exact))
