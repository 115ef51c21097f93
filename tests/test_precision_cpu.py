"""Low-precision Cholesky factors of C1's direct local systems as FP64
preconditioners, emulated on the CPU (numpy, FP32 arithmetic with every stored
entry rounded to FP32 or to BF16).  This pins the precision choice of the
engine's mixed direct solves: an FP32 factor takes a few refinement steps to
1e-11, a BF16 factor needs an order of magnitude more preconditioned
iterations (or breaks down on a non-positive pivot).

The systems are the explicit local matrices (N >= 256) of the oracle's C1 AMEn
run (2D Poisson 256 x 256, the reference's trajectory, one_site.py:49-75)."""

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import tt_oracle as O
from ttnumpy_amd import problems as P


def _bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(x.shape)


def _fp32(x):
    return np.asarray(x, dtype=np.float32)


def _chol(M, rnd):
    """Right-looking Cholesky in FP32 arithmetic, every stored entry rounded by rnd;
    None on a pivot that is not positive."""
    L = rnd(np.tril(M).astype(np.float32))
    for j in range(M.shape[0]):
        d = L[j, j]
        if not d > 0:
            return None
        d = np.float32(np.sqrt(d))
        L[j, j] = rnd(np.array([d]))[0]
        L[j + 1:, j] = rnd(L[j + 1:, j] / d)
        L[j + 1:, j + 1:] = rnd(L[j + 1:, j + 1:] - np.tril(np.outer(L[j + 1:, j], L[j + 1:, j])))
    return np.tril(L).astype(np.float64)


def _apply(L, r):
    return sl.solve_triangular(L.T, sl.solve_triangular(L, r, lower=True), lower=False)


def _refine(M, b, L, tol=1e-11, maxit=100):
    u, r, bn = np.zeros_like(b), b.copy(), np.linalg.norm(b)
    for it in range(1, maxit + 1):
        u += _apply(L, r)
        r = b - M @ u
        if np.linalg.norm(r) <= tol * bn:
            return it
    return None


def _pcg(M, b, L, tol=1e-11, maxit=500):
    x, r = np.zeros_like(b), b.copy()
    z = _apply(L, r)
    p, rz, bn = z.copy(), r @ z, np.linalg.norm(b)
    for it in range(1, maxit + 1):
        q = M @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(b - M @ x) <= tol * bn:
            return it
        z = _apply(L, r)
        rz, rz0 = r @ z, rz
        p = z + (rz / rz0) * p
    return None


@pytest.fixture(scope="module")
def c1_local_systems():
    mats = []
    orig = O.solve_local

    def capture(op, b_vec, iterative, *a, **k):
        if isinstance(op, np.ndarray) and op.shape[0] >= 256:
            mats.append((0.5 * (op + op.T), b_vec.copy()))
        return orig(op, b_vec, iterative, *a, **k)

    O.solve_local = capture
    try:
        O.amen(O.Train(P.poisson_2d_cores(8)), O.Train(P.ones_cores([2] * 16)),
               O.Train(P.random_train_cores([2] * 16, 2, 0)), symmetric=True, residual_tol=1e-8,
               max_sweeps=20, seed=0)
    finally:
        O.solve_local = orig
    mats.sort(key=lambda t: -t[0].shape[0])
    return mats


def test_fp32_and_bf16_factors_as_preconditioners(c1_local_systems):
    assert len(c1_local_systems) >= 40
    # the two largest systems and two of N = 800
    picks = c1_local_systems[:2] + [t for t in c1_local_systems if t[0].shape[0] == 800][:2]
    bf16_breakdowns, bf16_pcg = 0, []
    for M, b in picks:
        ev = np.linalg.eigvalsh(M)
        assert 1e4 < ev[-1] / ev[0] < 1e5  # kappa of the local systems
        L32 = _chol(M, _fp32)
        assert L32 is not None
        ir32, pcg32 = _refine(M, b, L32), _pcg(M, b, L32)
        assert ir32 is not None and ir32 <= 4 and pcg32 is not None and pcg32 <= 4, (ir32, pcg32)
        L16 = _chol(M, _bf16)
        if L16 is None:
            bf16_breakdowns += 1
            continue
        it = _pcg(M, b, L16)
        assert it is not None and it >= 10, it
        bf16_pcg.append(it)
    print("BF16 breakdowns", bf16_breakdowns, "BF16-preconditioned CG iterations", bf16_pcg)
    # measured: no breakdown in this emulation, 25, 12, 20 and 21 iterations
    assert bf16_breakdowns + len(bf16_pcg) == len(picks)
