"""Solver API behaviour of the device path against the oracle and the
contract of the reference's solver tests (tests/test_solvers_common.py,
tests/test_amen.py): in-place semantics, zero right-hand sides, rank caps,
threshold mode, default guesses, seeds, operand checks, the dense fallback."""

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle import tt_oracle as O  # noqa: E402
from ttnumpy_amd import problems as P  # noqa: E402

SOLVERS = ["amen", "als", "mals"]


def _tx():
    import ttnumpy_amd as tx
    return tx


def _lap1d(n_bits):
    tx = _tx()
    h = 1.0 / (2 ** n_bits + 1)
    A = tx.TensorTrain(P.laplacian_1d_cores(n_bits, h))
    b = tx.TensorTrain(P.ones_cores([2] * n_bits))
    return A, b


def _solver(name):
    return getattr(_tx(), name)


def _full(t):
    return np.asarray(t.get_full_tensor()).reshape(-1)


@pytest.mark.parametrize("name", SOLVERS)
def test_inplace_returns_the_guess_and_copy_leaves_it(name):
    print(f"api inplace {name}", flush=True)
    tx = _tx()
    solve = _solver(name)
    A, b = _lap1d(3)
    guess = tx.rand([2] * 3, ranks=2, seed=1)
    out = solve(A, b, x0=guess, inplace=True, max_sweeps=10, residual_tol=1e-8, symmetric=True)
    assert out is guess
    assert float(tx.relative_residual(A, b, guess)) <= 1e-8
    guess2 = tx.rand([2] * 3, ranks=2, seed=2)
    before = [c.copy() for c in guess2.to_numpy()]
    out2 = solve(A, b, x0=guess2, max_sweeps=10, residual_tol=1e-8, symmetric=True)
    assert out2 is not guess2
    for c0, c1 in zip(before, guess2.to_numpy()):
        np.testing.assert_array_equal(c0, c1)


@pytest.mark.parametrize("name", SOLVERS)
def test_zero_right_hand_side_gives_zero(name):
    print(f"api zero rhs {name}", flush=True)
    tx = _tx()
    A, _ = _lap1d(4)
    zero = tx.zeros([2] * 4)
    guess = tx.ones([2] * 4, ranks=2).orthonormalize(canonical="right")
    x, info = _solver(name)(A, zero, x0=guess, max_sweeps=2, residual_tol=1e-8, symmetric=True,
                            return_info=True)
    assert np.linalg.norm(_full(x)) == 0.0
    assert float(tx.relative_residual(A, zero, x)) == 0.0
    assert min(x.bonds) >= 1
    assert info.stop_reason is not None


@pytest.mark.parametrize("name", ["amen", "mals"])
def test_rank_cap_and_bad_caps(name):
    print(f"api {name} max_rank", flush=True)
    tx = _tx()
    A = tx.TensorTrain(P.poisson_2d_cores(3))
    b = tx.TensorTrain(P.ones_cores([2] * 6))
    kw = dict(seed=0) if name == "amen" else {}
    x = _solver(name)(A, b, max_rank=3, max_sweeps=4, residual_tol=1e-12, symmetric=True, **kw)
    assert max(x.bonds) <= 3
    for bad in (0, -2):
        with pytest.raises(ValueError):
            _solver(name)(A, b, max_rank=bad)


def test_rejects_a_ttm_right_hand_side():
    tx = _tx()
    A, _ = _lap1d(3)
    x0 = tx.ones([2] * 3)
    for name in SOLVERS:
        with pytest.raises(tx.WrongTTType, match="Right-hand side"):
            _solver(name)(A, A, x0=x0)


def test_amen_default_guess_and_threshold_mode_match_oracle():
    """x0=None starts from ones with b's bonds; without residual_tol the
    ranks follow the singular-value threshold (no round-off feedback through
    residual budgets), so bonds are compared exactly."""
    print("api amen default guess / threshold mode", flush=True)
    tx = _tx()
    A_c, b_c = P.poisson_2d_cores(3), P.ones_cores([2] * 6)
    x, info = tx.amen(tx.TensorTrain(A_c), tx.TensorTrain(b_c), max_sweeps=4, symmetric=True,
                      seed=0, return_info=True)
    xo, io = O.amen(O.Train(A_c), O.Train(b_c), max_sweeps=4, symmetric=True, seed=0)
    assert x.bonds == xo.bonds
    if max(info.residuals[-1], io.residuals[-1]) < 1e-12:
        # solved to round-off: the stagnation stop compares two rounding-level
        # residuals, so where it fires can move by one sweep
        assert abs(info.num_sweeps_run - io.num_sweeps_run) <= 1
    else:
        assert info.num_sweeps_run == io.num_sweeps_run
    ref = xo.full().reshape(-1)
    assert np.linalg.norm(_full(x) - ref) <= 1e-9 * np.linalg.norm(ref)


def test_amen_seed_reproducible():
    print("api amen seed", flush=True)
    tx = _tx()
    A = tx.TensorTrain(P.poisson_2d_cores(3))
    b = tx.TensorTrain(P.ones_cores([2] * 6))
    runs = [tx.amen(A, b, max_sweeps=3, residual_tol=1e-8, symmetric=True, seed=7) for _ in range(2)]
    for c0, c1 in zip(runs[0].to_numpy(), runs[1].to_numpy()):
        np.testing.assert_array_equal(c0, c1)


def test_dense_fallback_is_counted_when_armed():
    """A local tolerance MINRES cannot meet (1e-14), with allow_dense_fallback and
    every local system iterative, ends in the explicit-matrix solve and is
    counted as a fallback, not as inexact."""
    print("api dense fallback", flush=True)
    tx = _tx()
    A, b = _lap1d(6)
    x0 = tx.rand([2] * 6, ranks=2, seed=0)
    _, info = tx.amen(A, b, x0=x0, max_sweeps=1, residual_tol=1e-10, symmetric=True, seed=0,
                      local_solver="minres", direct_solve_size=0, local_rtol=1e-14,
                      allow_dense_fallback=True, stagnation_factor=None, return_info=True)
    assert info.local_solves > 0
    assert 0 < info.local_fallbacks <= info.local_solves
    assert info.local_inexact == 0


def _convection_diffusion_cores(n_bits, c):
    """-u'' + c u' on 2**n_bits points (Dirichlet, central differences): the QTT
    Laplacian plus c (S - S^T) / (2h) with S the up-shift, summed and rounded."""
    from ttnumpy_amd import diffusion3d as T3
    h = 1.0 / (2**n_bits + 1)
    S = T3.shift_cores(n_bits)
    St = [np.ascontiguousarray(x.transpose(0, 2, 1, 3)) for x in S]
    S = [x.copy() for x in S]
    St[0] = St[0] * (-1.0)
    S[0] = S[0] * (c / (2 * h))
    St[0] = St[0] * (c / (2 * h))
    lap = P.laplacian_1d_cores(n_bits, h)
    return P.round_cores(P.block_sum_cores(P.block_sum_cores(lap, S), St), 1e-14)


@pytest.mark.parametrize("name", SOLVERS)
def test_nonsymmetric_gmres_matches_oracle(name):
    """symmetric=False with iterative local solves (GMRES, restart 20) on a
    convection-diffusion operator: the device and the oracle (scipy's gmres,
    as the reference calls it) both reach 1e-8, in sweep counts within two,
    with solutions within the residual-implied distance of the dense solution.
    ALS starts from a rank-8 guess (the solution's ranks fit it)."""
    print(f"api {name} nonsymmetric gmres", flush=True)
    tx = _tx()
    L = 6
    A_c = _convection_diffusion_cores(L, 20.0)
    b_c = P.ones_cores([2] * L)
    kw = dict(residual_tol=1e-8, max_sweeps=10, symmetric=False, direct_solve_size=4)
    if name == "amen":
        kw.update(max_rank=16, seed=0)
    elif name == "mals":
        kw.update(max_rank=16)
    x0 = tx.rand([2] * L, ranks=8 if name == "als" else 2, seed=0)
    x, info = _solver(name)(tx.TensorTrain(A_c), tx.TensorTrain(b_c), x0=x0, return_info=True,
                            **kw)
    xo, io = getattr(O, name)(O.Train(A_c), O.Train(b_c), O.Train(x0.to_numpy()), **kw)
    Ad = O.Train(A_c).dense_matrix()
    assert np.linalg.norm(Ad - Ad.T) > 1.0        # really nonsymmetric
    ref = np.linalg.solve(Ad, np.ones(2**L))
    cond = np.linalg.cond(Ad)
    for stop, res in ((info.stop_reason, info.residuals), (io.stop_reason, io.residuals)):
        assert str(stop) in ("converged", "stagnated")
        assert res[-1] <= 3e-8
    assert info.local_solves > 0
    assert abs(info.num_sweeps_run - io.num_sweeps_run) <= 2
    for xf in (_full(x), xo.full().reshape(-1)):
        assert np.linalg.norm(xf - ref) <= 3e-8 * cond * np.linalg.norm(ref)
