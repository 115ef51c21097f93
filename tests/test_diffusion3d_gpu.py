"""The 3-D variable-coefficient diffusion builder and solve on the device:
exp(g) by scaling and squaring against the dense exponential, the operator
against a directly assembled dense finite-volume matrix, and AMEn on it
against the dense solution and the oracle."""

import numpy as np
import pytest

from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


def _dense_field(T3, spec, offsets, scale):
    C = T3.field_coefficients(spec)
    F = []
    for ax in range(3):
        t = T3._axis_points(spec.n_bits, offsets[ax])
        basis = np.stack([np.sin((m + 1) * np.pi * t) for m in range(spec.modes)])
        F.append(np.einsum("akb,kt->atb", C[ax], basis))
    return scale * np.einsum("axb,byc,czd->xyz", F[0], F[1], F[2])


def _dense_operator(T3, spec, scale):
    """sum_ax [diag(k+ + k-) - diag(k+) S - S^T diag(k+)] / h^2, k = exp(g) at the faces."""
    n = 2**spec.n_bits
    S1, I1 = np.eye(n, k=1), np.eye(n)
    A = np.zeros((n**3, n**3))
    for ax in range(3):
        facs = [I1, I1, I1]
        facs[ax] = S1
        S = np.kron(np.kron(facs[0], facs[1]), facs[2])
        off = [0.0, 0.0, 0.0]
        off[ax] = 0.5
        kp = np.exp(_dense_field(T3, spec, off, scale)).reshape(-1)
        off[ax] = -0.5
        km = np.exp(_dense_field(T3, spec, off, scale)).reshape(-1)
        A += np.diag(kp + km) - kp[:, None] * S - S.T * kp[None, :]
    return A * (n + 1) ** 2


@pytest.mark.parametrize("method", ["direct", "randomized"])
def test_exp_by_scaling_and_squaring(method):
    print(f"diffusion3d exp {method}", flush=True)
    from ttnumpy_amd import diffusion3d as T3
    import ttnumpy_amd as tx
    spec = T3.FieldSpec(n_bits=3, modes=2, seed=1)
    cores, scale = T3.field_cores(spec)
    bound = T3.field_bound(spec, scale)
    g = _dense_field(T3, spec, (0, 0, 0), scale)
    assert np.abs(g).max() <= bound
    k = T3.tt_exp(tx.TensorTrain(cores), bound, error=1e-13,
                  max_rank=64 if method == "randomized" else None, method=method)
    ref = np.exp(g).reshape(-1)
    err = np.linalg.norm(k.get_full_tensor().reshape(-1) - ref) / np.linalg.norm(ref)
    # s squarings double the relative rounding error of each product s times;
    # the sketched rounding adds its own (the sketch covers every rank here)
    assert err < (1e-10 if method == "direct" else 1e-9), err


def test_operator_matches_dense_finite_volumes():
    print("diffusion3d operator L=2,3", flush=True)
    from ttnumpy_amd import diffusion3d as T3
    for L in (2, 3):
        spec = T3.FieldSpec(n_bits=L, modes=2, seed=2)
        A, info = T3.operator(spec, k_error=1e-13, op_error=1e-13)
        _, scale = T3.field_cores(spec)
        n3 = 8**L
        dense = A.get_full_tensor().reshape(n3, n3)
        ref = _dense_operator(T3, spec, scale)
        assert np.linalg.norm(dense - ref) <= 1e-10 * np.linalg.norm(ref)
        assert np.linalg.norm(dense - dense.T) <= 1e-10 * np.linalg.norm(ref)
        assert np.linalg.eigvalsh(0.5 * (ref + ref.T)).min() > 0
        assert info.bonds == A.bonds and len(info.axis_bonds) == 3


def test_amen_on_variable_coefficient_3d_matches_dense_and_oracle():
    """L=3 (512 unknowns, 9 cores): AMEn reaches 1e-8 on the device and in the
    oracle, both solutions within the distance the residual allows of the
    dense solution."""
    print("diffusion3d amen L=3", flush=True)
    from threadpoolctl import threadpool_limits
    from ttnumpy_amd import diffusion3d as T3
    import ttnumpy_amd as tx
    spec = T3.FieldSpec(n_bits=3, modes=2, seed=0)
    A, _ = T3.operator(spec, k_error=1e-12, op_error=1e-13)
    b = T3.rhs(3)
    x0 = tx.rand([2] * 9, ranks=2, seed=0)
    kw = dict(max_rank=64, residual_tol=1e-8, max_sweeps=30, symmetric=True, seed=0)
    x, info = tx.amen(A, b, x0=x0, return_info=True, **kw)
    Ad = A.get_full_tensor().reshape(512, 512)
    ref = np.linalg.solve(Ad, np.ones(512))
    cond = np.linalg.cond(Ad)
    xd = x.get_full_tensor().reshape(-1)
    assert str(info.stop_reason) in ("converged", "stagnated")
    assert info.residuals[-1] <= 3e-8
    assert np.linalg.norm(xd - ref) <= 3e-8 * cond * np.linalg.norm(ref)
    with threadpool_limits(1):
        xo, io = O.amen(O.Train(A.to_numpy()), O.Train(b.to_numpy()),
                        O.Train(x0.to_numpy()), **kw)
    assert io.residuals[-1] <= 3e-8
    assert abs(info.num_sweeps_run - io.num_sweeps_run) <= 2
    assert np.linalg.norm(O.Train(xo.cores).full() - ref) <= 3e-8 * cond * np.linalg.norm(ref)


def test_amen_3d_minres_regime_matches_dense_and_oracle():
    """C3's iterative regime: L=4 (4096 unknowns, 12 cores) with
    direct_solve_size=64, so every local system above 64 unknowns takes the
    matrix-free MINRES path with the reference's restarts
    (one_site.py:144-162, common.py:776-809), as C3's systems of up to 8192
    unknowns do at full size.  Device and oracle converge, in sweep counts
    within two, with MINRES solves on both, and both solutions lie within the
    residual-implied distance of the dense solution."""
    print("diffusion3d amen L=4 minres", flush=True)
    from threadpoolctl import threadpool_limits
    from ttnumpy_amd import diffusion3d as T3
    import ttnumpy_amd as tx
    spec = T3.FieldSpec(n_bits=4, modes=2, seed=0)
    A, _ = T3.operator(spec, k_error=1e-12, op_error=1e-13)
    b = T3.rhs(4)
    x0 = tx.rand([2] * 12, ranks=2, seed=0)
    kw = dict(max_rank=64, residual_tol=1e-8, max_sweeps=30, symmetric=True, seed=0,
              direct_solve_size=64)
    x, info = tx.amen(A, b, x0=x0, return_info=True, **kw)
    Ad = A.get_full_tensor().reshape(4096, 4096)
    ref = np.linalg.solve(Ad, np.ones(4096))
    cond = np.linalg.cond(Ad)
    xd = x.get_full_tensor().reshape(-1)
    with threadpool_limits(8):
        xo, io = O.amen(O.Train(A.to_numpy()), O.Train(b.to_numpy()),
                        O.Train(x0.to_numpy()), **kw)
    print(info, io, cond, flush=True)
    assert info.local_solves > 0 and io.local_solves > 0
    for stop, res in ((info.stop_reason, info.residuals), (io.stop_reason, io.residuals)):
        assert str(stop) == "converged" and res[-1] <= 1e-8
    assert abs(info.num_sweeps_run - io.num_sweeps_run) <= 2
    for xf, res in ((xd, info.residuals[-1]), (O.Train(xo.cores).full().reshape(-1),
                                               io.residuals[-1])):
        assert np.linalg.norm(xf - ref) <= cond * res * np.linalg.norm(ref)
