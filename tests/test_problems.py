"""Input builders: the QTT operators against dense assemblies (CPU)."""

import numpy as np
import pytest

from ttnumpy_amd import problems as P


def _dense_ttm(cores):
    t = cores[0]
    for c in cores[1:]:
        t = np.tensordot(t, c, axes=(t.ndim - 1, 0))
    d = len(cores)
    n = [c.shape[1] for c in cores]
    t = t.reshape([x for k in range(d) for x in (n[k], cores[k].shape[2])])
    rows = [2 * k for k in range(d)]
    cols = [2 * k + 1 for k in range(d)]
    N = int(np.prod(n))
    return t.transpose(rows + cols).reshape(N, N)


def _diffusion_dense(L, theta):
    """Five-point finite-volume -div(k grad u), x index major, k at edge midpoints."""
    n = 2**L
    h = 1.0 / (n + 1)
    xs = (np.arange(n) + 1) * h
    xe = (np.arange(n + 1) + 0.5) * h

    def k(x, y):
        return 1.0 + theta * np.sin(np.pi * x) * np.sin(np.pi * y)

    A = np.zeros((n * n, n * n))
    for ix in range(n):
        for iy in range(n):
            r = ix * n + iy
            kxp, kxm = k(xe[ix + 1], xs[iy]), k(xe[ix], xs[iy])
            kyp, kym = k(xs[ix], xe[iy + 1]), k(xs[ix], xe[iy])
            A[r, r] = (kxp + kxm + kyp + kym) / h**2
            if ix + 1 < n:
                A[r, (ix + 1) * n + iy] = -kxp / h**2
            if ix > 0:
                A[r, (ix - 1) * n + iy] = -kxm / h**2
            if iy + 1 < n:
                A[r, ix * n + iy + 1] = -kyp / h**2
            if iy > 0:
                A[r, ix * n + iy - 1] = -kym / h**2
    return A


@pytest.mark.parametrize("L,theta", [(2, 0.0), (3, 0.7), (4, 0.9)])
def test_diffusion_2d_matches_dense_assembly(L, theta):
    A = _dense_ttm(P.diffusion_2d_cores(L, theta))
    ref = _diffusion_dense(L, theta)
    assert np.abs(A - ref).max() <= 1e-12 * np.abs(ref).max()


def test_diffusion_2d_theta_zero_is_poisson():
    L = 4
    a = _dense_ttm(P.diffusion_2d_cores(L, 0.0))
    b = _dense_ttm(P.poisson_2d_cores(L))
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_ttm_from_dense_round_trip_and_ranks():
    rng = np.random.default_rng(0)
    L = 5
    n = 2**L
    T = np.diag(rng.standard_normal(n - 1), 1)  # a shifted diagonal: QTT rank small
    cores = P.ttm_cores_from_dense(T, L)
    assert np.abs(_dense_ttm(cores) - T).max() <= 1e-13 * np.abs(T).max()
    s = np.sin(np.pi * (np.arange(n) + 1) / (n + 1))
    S = P.ttm_cores_from_dense(np.diag(s), L)
    assert max(c.shape[3] for c in S) == 2  # sin(a + b i) has QTT rank 2


def test_round_cores_keeps_the_operator():
    L = 4
    cores = P.block_sum_cores(P.poisson_2d_cores(L), P.poisson_2d_cores(L))  # 2 A, bonds doubled
    r = P.round_cores(cores, 1e-14)
    assert max(c.shape[3] for c in r) < max(c.shape[3] for c in cores)
    a = _dense_ttm(r)
    b = 2.0 * _dense_ttm(P.poisson_2d_cores(L))
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_shifted_laplacian_sum_small():
    d, n, c = 3, 4, 1.0 / 64
    A = _dense_ttm(P.shifted_laplacian_sum_cores(d, n, c))
    L1 = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1) + c * np.eye(n)
    I = np.eye(n)
    ref = (np.kron(np.kron(L1, I), I) + np.kron(np.kron(I, L1), I) + np.kron(np.kron(I, I), L1))
    assert np.abs(A - ref).max() <= 1e-13 * np.abs(ref).max()


def test_serving_stream_is_seeded():
    from ttnumpy_amd.serving import ServingConfig, request_stream
    a = request_stream(ServingConfig(n_requests=64, seed=3))
    b = request_stream(ServingConfig(n_requests=64, seed=3))
    assert [(x.level, x.theta, x.offset) for x in a] == [(x.level, x.theta, x.offset) for x in b]
    gaps = np.diff([x.offset for x in a])
    assert np.all(gaps >= 0) and 0.005 < gaps.mean() < 0.05  # rate 50/s
    assert {x.level for x in a} <= set(range(8, 13))
    assert all(0.0 <= x.theta < 0.9 for x in a)
