"""Kernel-level parity of libttx against NumPy FP64 on the same seeded inputs.

Every case prints its name and shape before it touches the device.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    from ttnumpy_amd import device as D
    return D


_KEEP = []


def _t(a):
    """Device copy that stays alive for the whole module (kernels read it after enqueue)."""
    t = _dev().to_device(a)
    _KEEP.append(t)
    return t


def _h(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("M,N,K,batch", [(1, 1, 1, 1), (5, 7, 3, 1), (64, 64, 16, 1), (65, 129, 33, 3),
                                         (200, 17, 300, 2), (3, 500, 7, 5)])
@pytest.mark.parametrize("layout", ["nn", "tn", "nt", "tt"])
def test_gemm_strided(M, N, K, batch, layout):
    print(f"gemm M={M} N={N} K={K} batch={batch} layout={layout}", flush=True)
    D = _dev()
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A = rng.standard_normal((batch, M, K)) if layout[0] == "n" else rng.standard_normal((batch, K, M))
    B = rng.standard_normal((batch, K, N)) if layout[1] == "n" else rng.standard_normal((batch, N, K))
    C0 = rng.standard_normal((batch, M, N))
    Ad, Bd, Cd = _t(A), _t(B), _t(C0)
    a_ms, a_ks = (K, 1) if layout[0] == "n" else (1, M)
    b_ks, b_ns = (N, 1) if layout[1] == "n" else (1, K)
    D.gemm(batch, M, N, K, 1.5, D.P(Ad), a_ms, a_ks, M * K, D.P(Bd), b_ks, b_ns, K * N, -0.5,
           D.P(Cd), N, 1, M * N)
    Am = A if layout[0] == "n" else A.transpose(0, 2, 1)
    Bm = B if layout[1] == "n" else B.transpose(0, 2, 1)
    ref = 1.5 * Am @ Bm - 0.5 * C0
    np.testing.assert_allclose(_h(Cd), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_gemm_tilings_agree_bitwise():
    """A product large enough for the 64 x 64 tiling (525 tiles) and one row
    block of it, which runs on the 32 x 32 tiling, give identical values."""
    print("gemm 64x64 vs 32x32 tiling", flush=True)
    D = _dev()
    rng = np.random.default_rng(11)
    M, N, K = 1600, 1344, 40
    A, B = rng.standard_normal((M, K)), rng.standard_normal((K, N))
    Ad, Bd = _t(A), _t(B)
    C = D.empty(M, N)
    D.gemm(1, M, N, K, 1.0, D.P(Ad), K, 1, 0, D.P(Bd), N, 1, 0, 0.0, D.P(C), N, 1, 0)
    part = D.empty(320, N)
    D.gemm(1, 320, N, K, 1.0, D.P(Ad, 320 * K), K, 1, 0, D.P(Bd), N, 1, 0, 0.0, D.P(part), N, 1, 0)
    full = _h(C)
    np.testing.assert_allclose(full, A @ B, rtol=1e-12, atol=1e-12 * np.abs(full).max())
    np.testing.assert_array_equal(_h(part), full[320:640])


def test_gemm_taller_than_one_launch():
    """More rows than one launch's grid holds (65535 row tiles): the product
    runs over row blocks and equals the reference everywhere, the last
    partial block included."""
    print("gemm tall M=2.2e6", flush=True)
    D = _dev()
    rng = np.random.default_rng(3)
    M, N, K = 2_200_003, 3, 5
    A, B = rng.standard_normal((M, K)), rng.standard_normal((K, N))
    C = D.empty(M, N)
    D.gemm(1, M, N, K, 1.0, D.P(_t(A)), K, 1, 0, D.P(_t(B)), N, 1, 0, 0.0, D.P(C), N, 1, 0)
    np.testing.assert_allclose(_h(C), A @ B, rtol=1e-12, atol=1e-12)


def test_gemm_beta_zero_ignores_nan():
    D = _dev()
    A = np.ones((4, 3))
    B = np.ones((3, 5))
    C = np.full((4, 5), np.nan)
    Cd = _t(C)
    D.gemm(1, 4, 5, 3, 1.0, D.P(_t(A)), 3, 1, 0, D.P(_t(B)), 5, 1, 0, 0.0, D.P(Cd), 5, 1, 0)
    np.testing.assert_array_equal(_h(Cd), np.full((4, 5), 3.0))


def test_copy_nd_permute_and_norm():
    D = _dev()
    rng = np.random.default_rng(1)
    X = rng.standard_normal((3, 4, 5))
    Xd = _t(X)
    Y = D.empty(5, 3, 4)
    D.copy_nd([3, 4, 5], D.P(Xd), [20, 5, 1], D.P(Y), [4, 1, 12], 2.0)
    np.testing.assert_array_equal(_h(Y), 2.0 * X.transpose(2, 0, 1))
    assert float(D.norm(Xd).item()) == pytest.approx(np.linalg.norm(X), rel=1e-14)
    big = rng.standard_normal(300001)
    assert float(D.norm(_t(big)).item()) == pytest.approx(np.linalg.norm(big), rel=1e-13)


@pytest.mark.parametrize("m,n", [(1, 1), (6, 3), (3, 6), (64, 64), (130, 33), (33, 130), (512, 128),
                                 (257, 200)])
def test_qr(m, n):
    print(f"qr {m}x{n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(m + 7 * n)
    A = rng.standard_normal((m, n))
    k = min(m, n)
    Q = D.empty(m, k)
    R = D.empty(k, n)
    D.qr(m, n, D.P(_t(A)), n, 1, D.P(Q), k, 1, D.P(R), n, 1)
    Qh, Rh = _h(Q), _h(R)
    np.testing.assert_allclose(Qh @ Rh, A, atol=1e-13 * np.abs(A).max() * max(m, n))
    np.testing.assert_allclose(Qh.T @ Qh, np.eye(k), atol=1e-13 * max(m, n))
    assert np.allclose(np.tril(Rh, -1), 0.0)
    # gauge-invariant: |diag R| equals LAPACK's
    Rref = np.linalg.qr(A, mode="reduced")[1]
    np.testing.assert_allclose(np.abs(np.diag(Rh)), np.abs(np.diag(Rref)), rtol=1e-11)


def test_qr_rank_deficient_and_transposed_view():
    D = _dev()
    rng = np.random.default_rng(5)
    A = rng.standard_normal((40, 3)) @ rng.standard_normal((3, 10))  # rank 3
    A[:, 4] = 0.0
    Ad = _t(A)
    # QR of A^T read through strides (the right-sweep unfolding)
    Q = D.empty(10, 10)
    R = D.empty(10, 40)
    D.qr(10, 40, D.P(Ad), 1, 10, D.P(Q), 10, 1, D.P(R), 40, 1)
    np.testing.assert_allclose(_h(Q) @ _h(R), A.T, atol=1e-12)
    np.testing.assert_allclose(_h(Q).T @ _h(Q), np.eye(10), atol=1e-12)


@pytest.mark.parametrize("case", ["kappa1e4", "kappa1e9", "rank_deficient", "transposed_view"])
def test_qr_tall_cholesky_qr2_and_fallback(case):
    """Tall QRs with Q (m >= 256, 64 <= n <= 128) run CholeskyQR2 on the MFMA
    GEMM; inputs too ill-conditioned for two passes (kappa ~1e9), or rank
    deficient, are refused on the device and refactored by Householder.  Every
    case must give an orthonormal Q and QR = A to rounding."""
    print(f"qr tall {case}", flush=True)
    D = _dev()
    rng = np.random.default_rng(11)
    m, n = 512, 128
    A = rng.standard_normal((m, n))
    if case == "kappa1e4":
        A = A * np.logspace(0, -4, n)[None, :]
    elif case == "kappa1e9":
        A = A * np.logspace(0, -9, n)[None, :]
    elif case == "rank_deficient":
        A[:, 100:] = A[:, :28]
    Q = D.empty(m, n)
    R = D.empty(n, n)
    if case == "transposed_view":
        At = _t(np.ascontiguousarray(A.T))  # n x m storage, read as its transpose
        D.qr(m, n, D.P(At), 1, m, D.P(Q), n, 1, D.P(R), n, 1)
    else:
        D.qr(m, n, D.P(_t(A)), n, 1, D.P(Q), n, 1, D.P(R), n, 1)
    Qh, Rh = _h(Q), _h(R)
    np.testing.assert_allclose(Qh @ Rh, A, atol=1e-13 * np.abs(A).max() * m)
    np.testing.assert_allclose(Qh.T @ Qh, np.eye(n), atol=1e-13 * m)
    assert np.allclose(np.tril(Rh, -1), 0.0)
    if case != "rank_deficient":
        Rref = np.linalg.qr(A, mode="reduced")[1]
        np.testing.assert_allclose(np.abs(np.diag(Rh)), np.abs(np.diag(Rref)), rtol=1e-9)


@pytest.mark.parametrize("m,n", [(384, 128), (128, 384), (512, 96), (96, 300)])
def test_svd_tall_cholesky_qr2_path(m, n):
    """Large tall/wide SVDs: CholeskyQR2 of the long side, the Jacobi kernel on
    the square factor without accumulating rotations, V from the inverse
    Cholesky factors, accepted by the orthogonality certificate (kappa ~ 30
    here; the kappa ~ 1e9 cases of test_svd take the fallbacks)."""
    print(f"svd tall {m}x{n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(m + n)
    k = min(m, n)
    A = rng.standard_normal((m, n))
    if m >= n:
        A = A * np.logspace(0, -1.5, n)[None, :]
    else:
        A = np.logspace(0, -1.5, m)[:, None] * A
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    U, S, Vt = D.svd(m, n, D.P(_t(A)), n, 1, info)
    assert int(info[0].item()) == 0
    Uh, Sh, Vh = _h(U), _h(S), _h(Vt)
    Sref = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(Sh, Sref, rtol=1e-12)
    np.testing.assert_allclose((Uh * Sh) @ Vh, A, atol=1e-13 * Sref[0] * max(m, n))
    np.testing.assert_allclose(Uh.T @ Uh, np.eye(k), atol=1e-13 * max(m, n))
    np.testing.assert_allclose(Vh @ Vh.T, np.eye(k), atol=1e-13 * max(m, n))


@pytest.mark.parametrize("m,n", [(384, 128), (256, 64), (128, 384)])
def test_svd_clustered_spectrum_orthonormal(m, n):
    """Clustered singular values (X = orthonormal * (I + 1e-8 E)): the rotations
    of such pairs keep O(1) angles while |c| is small, so the V-free Jacobi of
    the CholeskyQR2 path must not stop on |c| alone; U and V stay orthonormal
    to the rounding level and the singular values match LAPACK's."""
    print(f"svd clustered {m}x{n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(11 * m + n)
    k = min(m, n)
    Ux = np.linalg.qr(rng.standard_normal((m, k)))[0]
    Vx = np.linalg.qr(rng.standard_normal((n, k)))[0]
    A = Ux @ (np.eye(k) + 1e-8 * rng.standard_normal((k, k))) @ Vx.T
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    U, S, Vt = D.svd(m, n, D.P(_t(A)), n, 1, info)
    assert int(info[0].item()) == 0
    Uh, Sh, Vh = _h(U), _h(S), _h(Vt)
    np.testing.assert_allclose(Sh, np.linalg.svd(A, compute_uv=False), rtol=1e-13)
    np.testing.assert_allclose((Uh * Sh) @ Vh, A, atol=1e-13 * max(m, n))
    assert np.max(np.abs(Uh.T @ Uh - np.eye(k))) <= 1e-13 * max(m, n)
    assert np.max(np.abs(Vh @ Vh.T - np.eye(k))) <= 1e-13 * max(m, n)


@pytest.mark.parametrize("m,n", [(1, 1), (1, 5), (5, 1), (8, 8), (66, 33), (33, 66), (128, 384),
                                 (384, 128), (200, 150), (2, 2)])
def test_svd(m, n):
    print(f"svd {m}x{n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(3 * m + n)
    A = rng.standard_normal((m, n)) * np.logspace(0, -9, n)[None, :]
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    U, S, Vt = D.svd(m, n, D.P(_t(A)), n, 1, info)
    assert int(info[0].item()) == 0
    Uh, Sh, Vh = _h(U), _h(S), _h(Vt)
    Sref = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(Sh, Sref, rtol=1e-11, atol=1e-14 * Sref[0])
    assert np.all(np.diff(Sh) <= 0)
    np.testing.assert_allclose((Uh * Sh) @ Vh, A, atol=1e-13 * Sref[0] * max(m, n))
    k = min(m, n)
    np.testing.assert_allclose(Uh.T @ Uh, np.eye(k), atol=1e-12 * max(m, n))
    np.testing.assert_allclose(Vh @ Vh.T, np.eye(k), atol=1e-12 * max(m, n))


def test_svd_nonfinite_reports():
    D = _dev()
    A = np.ones((4, 3))
    A[1, 2] = np.nan
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    D.svd(4, 3, D.P(_t(A)), 3, 1, info)
    assert int(info[0].item()) == 2


@pytest.mark.parametrize("n", [1, 2, 17, 64, 100, 333, 800, 1000])
def test_lu_solve(n):
    print(f"lu n={n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    b = rng.standard_normal(n)
    Mb = np.concatenate([M, b[:, None]], axis=1).reshape(-1)  # row-major [M | b]
    u = D.empty(n)
    info = torch.zeros(1, dtype=torch.int32, device=D.device())
    D.lu_solve(_t(Mb), n, u, info)
    ref = np.linalg.solve(M, b)
    assert int(info.item()) == 0
    np.testing.assert_allclose(_h(u), ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("n,kind", [(129, "gauss"), (300, "gauss"), (1000, "gauss"),
                                    (257, "reversed"), (700, "reversed")])
def test_lu_solve_pivots_across_blocks(n, kind):
    """Pivots that come from far below the current block (the 128-column outer
    panels and 16-column sub-panels must see the same row order)."""
    print(f"lu pivoting n={n} {kind}", flush=True)
    D = _dev()
    rng = np.random.default_rng(1000 + n)
    if kind == "gauss":
        M = rng.standard_normal((n, n))  # partial pivoting swaps most rows
    else:
        # diagonally dominant rows stored bottom-up: every pivot lies in the last rows
        M = (rng.standard_normal((n, n)) + 4.0 * n ** 0.5 * np.eye(n))[::-1].copy()
    b = rng.standard_normal(n)
    ld = n + 3  # a padded row stride
    Mb = np.zeros((n, ld))
    Mb[:, :n] = M
    Mb[:, n] = b
    u = D.empty(n)
    info = torch.zeros(1, dtype=torch.int32, device=D.device())
    D.lu_solve(_t(Mb.reshape(-1)), n, u, info, ld=ld)
    assert int(info.item()) == 0
    x = _h(u)
    ref = np.linalg.solve(M, b)
    assert np.linalg.norm(M @ x - b) <= 1e-12 * np.linalg.norm(M, 1) * np.linalg.norm(x)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 100, 129, 300, 1000])
def test_chol_solve_spd(n):
    """Cholesky on the lower triangle; the upper triangle holds garbage on purpose."""
    print(f"chol n={n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(2000 + n)
    G = rng.standard_normal((n, n))
    M = G @ G.T / n + 0.5 * np.eye(n)  # SPD, condition ~ 1e1..1e3
    b = rng.standard_normal(n)
    ld = n + 2
    Mb = np.zeros((n, ld))
    Mb[:, :n] = np.tril(M) + np.triu(np.full((n, n), 7.0), 1)
    Mb[:, n] = b
    u = D.empty(n)
    info = torch.zeros(1, dtype=torch.int32, device=D.device())
    D.chol_solve(_t(Mb.reshape(-1)), n, u, info, ld=ld)
    assert int(info.item()) == 0
    x = _h(u)
    ref = np.linalg.solve(M, b)
    assert np.linalg.norm(M @ x - b) <= 1e-13 * np.linalg.norm(M, 1) * np.linalg.norm(x) * n ** 0.5
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("n,where", [(50, 0), (300, 137), (700, 699)])
def test_chol_reports_not_spd(n, where):
    """An indefinite symmetric matrix: the pivot at `where` turns negative."""
    D = _dev()
    rng = np.random.default_rng(3000 + n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.linspace(1.0, 2.0, n)
    M = (Q * lam) @ Q.T
    # make the leading (where+1) x (where+1) block indefinite at its last pivot
    M[where, where] -= 10.0 * n
    Mb = np.concatenate([M, np.ones((n, 1))], axis=1).reshape(-1)
    u = D.empty(n)
    info = torch.zeros(1, dtype=torch.int32, device=D.device())
    D.chol_solve(_t(Mb), n, u, info)
    assert int(info.item()) == 6  # TTX_ENOTSPD


def test_lu_singular_reports():
    D = _dev()
    n = 200
    rng = np.random.default_rng(7)
    M = rng.standard_normal((n, n))
    M[:, 150] = 0.0  # an exactly zero column: the pivot of column 150 is 0
    Mb = np.concatenate([M, np.ones((n, 1))], axis=1).reshape(-1)
    u = D.empty(n)
    info = torch.zeros(1, dtype=torch.int32, device=D.device())
    D.lu_solve(_t(Mb), n, u, info)
    assert int(info.item()) == 5  # TTX_ESINGULAR


def test_lu_needs_pivoting():
    D = _dev()
    M = np.array([[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [4.0, 5.0, 0.0]])
    b = np.array([1.0, 2.0, 3.0])
    Mb = np.concatenate([M, b[:, None]], axis=1).reshape(-1)
    u = D.empty(3)
    info = torch.zeros(1, dtype=torch.int32, device=D.device())
    D.lu_solve(_t(Mb), 3, u, info)
    np.testing.assert_allclose(_h(u), np.linalg.solve(M, b), rtol=1e-13)


@pytest.mark.parametrize("kappa,tight", [(10.0, True), (1e3, False)])
def test_minres_matches_scipy(kappa, tight):
    """Same stop iteration as SciPy's MINRES; iterates equal where Lanczos keeps orthogonality.

    At kappa = 1e3 the Lanczos vectors of two FP64 implementations lose
    orthogonality differently after ~sqrt(kappa) steps, so there the check is
    the iteration count and the true residual rather than the iterate.
    """
    import scipy.sparse.linalg as spla
    from ttnumpy_amd.solvers.common import minres_device
    D = _dev()
    rng = np.random.default_rng(0)
    n = 200
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    M = (Q * np.logspace(0, np.log10(kappa), n)) @ Q.T
    M = 0.5 * (M + M.T)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    op = D.ExplicitOp(_t(M.reshape(-1)), n)  # row-major storage
    for rtol in (1e-4, 1e-8, 1e-10):
        its = []
        ref, _ = spla.minres(M, b, x0=x0, rtol=rtol, callback=lambda xk: its.append(1))
        st = []
        got = _h(minres_device(op, _t(b), _t(x0), rtol, state_out=st))
        print(f"kappa {kappa} rtol {rtol}: scipy iterations {len(its)}, device iterations "
              f"{int(st[0][1])} istop {int(st[0][0])}", flush=True)
        assert abs(int(st[0][1]) - len(its)) <= (0 if tight else 1)
        r_ref = np.linalg.norm(M @ ref - b)
        r_got = np.linalg.norm(M @ got - b)
        assert r_got <= 1.5 * r_ref + 1e-14 * np.linalg.norm(b)
        if tight:
            np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())


def test_truncated_bond_kernel_known_answers():
    from oracle import tt_oracle as O
    D = _dev()
    out = torch.zeros(2, dtype=torch.int32, device=D.device())
    cases = [([4.0, 2.0, 1.0, 0.5], 1.0, None), ([1.0] + [0.099] * 100, 0.1, None),
             ([1.0, 0.05], 0.1, None), ([4.0, 2.0, 1.0, 0.5], 100.0, None),
             ([4.0, 2.0, 1.0, 0.5], 0.0, 2), ([1.0, 1e-17, 3e-18], 0.0, None),
             ([1e8, 1e-9], 0.0, None), ([1e8, 1e-5], 0.0, None), ([1.0, 1e-9], 0.0, None)]
    for s, delta, mr in cases:
        S = _t(np.array(s))
        one = _t(np.array([1.0]))
        D.truncated_bond_dev(S, one, delta, mr, out, 1)
        assert int(out[1].item()) == O.truncated_bond(np.array(s), delta, mr), (s, delta, mr)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (37, 65, 129), (200, 96, 512)])
def test_gemm_f32acc_matches_fp32_product(M, N, K):
    """FP32 contraction of FP64 arrays: equals the product of the FP32-rounded
    operands to FP32 accumulation error, and the FP64 product to FP32 error."""
    print(f"gemm f32acc {M}x{N}x{K}", flush=True)
    D = _dev()
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C = D.empty(M * N)
    from ttnumpy_amd import _lib
    rc = D.lib().ttx_gemm_f32acc(D.stream(), 1, M, N, K, 1.0, D.P(_t(A.reshape(-1))), K, 1, 0,
                                 D.P(_t(B.reshape(-1))), N, 1, 0, 0.0, D.P(C), N, 1, 0, None, None)
    _lib.check(rc, "ttx_gemm_f32acc")
    got = _h(C).reshape(M, N)
    ref32 = A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)
    scale = np.abs(A) @ np.abs(B)
    assert np.all(np.abs(got - ref32) <= 2 * K * 6e-8 * scale + 1e-30)
    assert np.abs(got - A @ B).max() <= 4 * K * 6e-8 * scale.max()
    # with device scales (operands far outside FP32's range) the result is the same
    big = 1e45
    sc = _t(np.array([np.linalg.norm(A) * big, np.linalg.norm(B)]))
    rc = D.lib().ttx_gemm_f32acc(D.stream(), 1, M, N, K, 1.0, D.P(_t((A * big).reshape(-1))), K, 1,
                                 0, D.P(_t(B.reshape(-1))), N, 1, 0, 0.0, D.P(C), N, 1, 0,
                                 D.P(sc), D.P(sc, 1))
    _lib.check(rc, "ttx_gemm_f32acc")
    got2 = _h(C).reshape(M, N) / big
    assert np.all(np.isfinite(got2))
    assert np.abs(got2 - A @ B).max() <= 4 * K * 6e-8 * scale.max()


def test_autotune_sets_a_limit_and_keeps_results():
    """Tuning the GEMM tiling on a workload changes its speed only: the
    truncation gives the same bits before and after."""
    print("gemm tiling autotune", flush=True)
    import ttnumpy_amd as tx
    from ttnumpy_amd import autotune as AT
    from ttnumpy_amd import problems as Pb
    x = tx.TensorTrain(Pb.random_train_cores([4] * 10, 24, 1, normal=True))
    before = x.truncate(error=1e-10, max_rank=16).to_numpy()
    old = AT.get_limit()
    try:
        rep = AT.tune(lambda: x.truncate(error=1e-10, max_rank=16), reps=3)
        assert rep["distinct_shapes"] > 0 and AT.get_limit() == rep["limit"]
        assert rep["predicted_ms"]["tuned"] <= rep["predicted_ms"]["previous"] + 1e-9
        after = x.truncate(error=1e-10, max_rank=16).to_numpy()
        for a, b in zip(before, after):
            np.testing.assert_array_equal(a, b)
    finally:
        AT.set_limit(old)


@pytest.mark.parametrize("n,seed,with_x0", [(60, 1, False), (300, 2, False), (300, 3, True),
                                            (45, 4, True)])
def test_gmres_matches_scipy(n, seed, with_x0):
    """Restarted GMRES on dense nonsymmetric systems (restart 20, so several
    cycles run) against scipy.sparse.linalg.gmres with the same rtol and x0:
    both meet the tolerance and agree to the distance it allows."""
    print(f"gmres n={n} seed={seed} x0={with_x0}", flush=True)
    import scipy.sparse.linalg as spla
    from ttnumpy_amd.solvers.common import gmres_device
    D = _dev()
    rng = np.random.default_rng(seed)
    # shift 1.5: 43-56 inner steps, so three restart cycles of 20 run
    A = np.eye(n) * 1.5 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n) * 0.1 if with_x0 else None
    rtol = 1e-10
    ref, info = spla.gmres(A, b, x0=x0, rtol=rtol, atol=0.0)
    assert info == 0
    steps = []
    spla.gmres(A, b, x0=x0, rtol=rtol, atol=0.0, callback=lambda pr: steps.append(pr),
               callback_type="pr_norm")
    op = D.ExplicitOp(_t(A.copy()), n)
    st = []
    x = _h(gmres_device(op, _t(b), _t(x0) if x0 is not None else None, rtol, state_out=st))
    assert np.linalg.norm(A @ x - b) <= rtol * np.linalg.norm(b) * 1.0001
    # the same iteration: the same number of inner (Arnoldi) steps as scipy's
    assert abs(int(st[0][13]) - len(steps)) <= 1, (int(st[0][13]), len(steps))
    cond = np.linalg.cond(A)
    assert np.linalg.norm(x - ref) <= 2 * cond * rtol * np.linalg.norm(ref)


def test_gmres_special_cases():
    """b = 0 returns zeros; an exact x0 returns at once; the identity breaks
    down after one step with the exact solution (scipy's early exits)."""
    print("gmres special cases", flush=True)
    from ttnumpy_amd.solvers.common import gmres_device
    D = _dev()
    n = 30
    rng = np.random.default_rng(9)
    A = np.eye(n) * 3.0 + rng.standard_normal((n, n)) * 0.1
    op = D.ExplicitOp(_t(A.copy()), n)
    np.testing.assert_array_equal(_h(gmres_device(op, _t(np.zeros(n)), None, 1e-8)), np.zeros(n))
    xs = rng.standard_normal(n)
    b = A @ xs
    x = _h(gmres_device(op, _t(b), _t(xs), 1e-8))
    assert np.linalg.norm(A @ x - b) <= 1e-8 * np.linalg.norm(b)
    eye = D.ExplicitOp(_t(np.eye(n)), n)
    x = _h(gmres_device(eye, _t(b), None, 1e-12))
    np.testing.assert_allclose(x, b, rtol=1e-14)


def test_gmres_local_cycle_matches_explicit_operator():
    """The cycle issued from C on the matrix-free one-site operator and the
    host-driven cycle on its explicit matrix solve the same (nonsymmetric,
    well-conditioned) system to the tolerance."""
    print("gmres local cycle vs explicit", flush=True)
    from ttnumpy_amd.solvers.common import gmres_device
    D = _dev()
    rng = np.random.default_rng(12)
    r, n_ = 4, 5
    psi = np.zeros((r, 2, r))
    phi = np.zeros((r, 2, r))
    A = np.zeros((2, n_, n_, 2))
    psi[:, 0, :] = np.eye(r)
    phi[:, 0, :] = np.eye(r)
    psi[:, 1, :] = 0.3 * rng.standard_normal((r, r))
    phi[:, 1, :] = 0.3 * rng.standard_normal((r, r))
    A[0, :, :, 0] = 4.0 * np.eye(n_) + rng.standard_normal((n_, n_)) / np.sqrt(n_)
    A[1, :, :, 1] = 0.3 * rng.standard_normal((n_, n_))
    op = D.LocalOp(_t(psi), _t(A), _t(phi))
    N = op.rows
    M = D.empty(N * N)
    op.explicit(M, N)
    Mh = _h(M).reshape(N, N)
    b = rng.standard_normal(N)
    rtol = 1e-11
    x_ex = _h(gmres_device(D.ExplicitOp(_t(Mh.copy()), N), _t(b), None, rtol))
    x_lo = _h(gmres_device(op, _t(b), None, rtol))
    for x in (x_ex, x_lo):
        assert np.linalg.norm(Mh @ x - b) <= 1.01 * rtol * np.linalg.norm(b)
    assert np.linalg.norm(x_lo - x_ex) <= 2 * np.linalg.cond(Mh) * rtol * np.linalg.norm(x_ex)


@pytest.mark.parametrize("m,n,zero_cols", [(1500, 700, False), (9000, 128, False),
                                            (4482, 737, False), (1200, 1200, False),
                                            (1842, 265, False), (512, 1181, False),
                                            (64, 7169, False), (2000, 200, True),
                                            (300, 2000, True)])
def test_qr_r_only_blocked_large(m, n, zero_cols):
    """R without Q of large matrices (the blocked path across the chip, tall
    and wide, with exactly zero columns for reflectors with tau = 0): R^T R
    equals A^T A, R is upper triangular, and |diag R| equals LAPACK's."""
    print(f"qr R-only blocked {m}x{n} zero_cols={zero_cols}", flush=True)
    D = _dev()
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    if zero_cols:
        A[:, 40:75] = 0.0
    k = min(m, n)
    Ad = _t(A)
    R = D.empty(k, n)
    D.qr(m, n, D.P(Ad), n, 1, None, 1, n, D.P(R), n, 1)
    Rh = _h(R)
    assert np.allclose(np.tril(Rh, -1), 0.0)
    scale = np.linalg.norm(A) ** 2
    np.testing.assert_allclose(Rh.T @ Rh, A.T @ A, rtol=0, atol=1e-13 * scale)
    ref = np.linalg.qr(A, mode="r")
    np.testing.assert_allclose(np.abs(np.diag(Rh)), np.abs(np.diag(ref)), rtol=1e-10,
                               atol=1e-12 * np.abs(np.diag(ref)).max())


@pytest.mark.parametrize("m,n,case", [(1500, 300, "graded"), (300, 1500, "graded"),
                                      (800, 200, "deficient"), (200, 800, "deficient"),
                                      (600, 150, "gauss"), (500, 135, "graded"), (700, 171, "gauss"),
                                      (171, 700, "deficient"), (255, 255, "graded")])
def test_svd_large_preconditioned_jacobi(m, n, case):
    """The large SVDs (columns taken by decreasing norm, Jacobi on the columns
    of R^T; on one workgroup below 139 columns, one workgroup per column pair
    and one launch per round from 139 on, odd counts with a dummy column):
    singular values to LAPACK's, the reconstruction, and
    orthonormal singular vectors for every singular value above the rounding
    level; exactly zero columns and repeated columns (ties in the norm order,
    singular values that are zero) included."""
    print(f"svd preconditioned {m}x{n} {case}", flush=True)
    D = _dev()
    rng = np.random.default_rng(m + 3 * n)
    k = min(m, n)
    if case == "graded":
        Ux = np.linalg.qr(rng.standard_normal((m, k)))[0]
        Vx = np.linalg.qr(rng.standard_normal((n, k)))[0]
        A = (Ux * np.logspace(0, -12, k)) @ Vx.T
    else:
        A = rng.standard_normal((m, n))
        if case == "deficient":
            if m >= n:
                A[:, 20:40] = 0.0
                A[:, 100:110] = A[:, 0:10]
            else:
                A[20:40, :] = 0.0
                A[100:110, :] = A[0:10, :]
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    U, S, Vt = D.svd(m, n, D.P(_t(A)), n, 1, info)
    assert int(info[0].item()) == 0
    Uh, Sh, Vh = _h(U), _h(S), _h(Vt)
    Sref = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(Sh, Sref, rtol=1e-10, atol=1e-13 * Sref[0])
    assert np.all(np.diff(Sh) <= 0)
    np.testing.assert_allclose((Uh * Sh) @ Vh, A, atol=1e-13 * Sref[0] * max(m, n))
    live = Sh > 1e-10 * Sh[0]
    kl = int(live.sum())
    assert kl == int((Sref > 1e-10 * Sref[0]).sum())
    if case == "deficient":
        assert kl == k - 30
    np.testing.assert_allclose(Uh[:, live].T @ Uh[:, live], np.eye(kl), atol=1e-12 * max(m, n))
    np.testing.assert_allclose(Vh[live] @ Vh[live].T, np.eye(kl), atol=1e-12 * max(m, n))
    # every singular vector, those of zero singular values included (LAPACK's are orthonormal)
    np.testing.assert_allclose(Uh.T @ Uh, np.eye(k), atol=1e-12 * max(m, n))
    np.testing.assert_allclose(Vh @ Vh.T, np.eye(k), atol=1e-12 * max(m, n))


@pytest.mark.parametrize("m,n", [(40, 12), (12, 40), (384, 128), (128, 384), (300, 160),
                                 (160, 300), (7, 7)])
def test_svd_exact_zero_singular_values_orthonormal(m, n):
    """Exactly zero singular values (zero columns or rows): the singular vectors
    that go with them are completed to orthonormal sets, as LAPACK's are, on
    every path (one workgroup, CholeskyQR2 with its Householder fallback,
    multi-workgroup Jacobi, wide and tall)."""
    print(f"svd zero singular values {m}x{n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(5 * m + n)
    k = min(m, n)
    A = rng.standard_normal((m, n))
    if m >= n:
        A[:, 3:7] = 0.0
    else:
        A[3:7, :] = 0.0
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    U, S, Vt = D.svd(m, n, D.P(_t(A)), n, 1, info)
    assert int(info[0].item()) == 0
    Uh, Sh, Vh = _h(U), _h(S), _h(Vt)
    assert np.count_nonzero(Sh == 0.0) == 4
    np.testing.assert_allclose(Sh, np.linalg.svd(A, compute_uv=False), rtol=1e-11, atol=1e-13 * Sh[0])
    np.testing.assert_allclose((Uh * Sh) @ Vh, A, atol=1e-13 * Sh[0] * max(m, n))
    np.testing.assert_allclose(Uh.T @ Uh, np.eye(k), atol=1e-12 * max(m, n))
    np.testing.assert_allclose(Vh @ Vh.T, np.eye(k), atol=1e-12 * max(m, n))


def test_svd_large_nonfinite_reports():
    """A NaN in a factor of the multi-workgroup Jacobi path is reported as on
    the small path (status 2), without running the sweeps."""
    D = _dev()
    A = np.random.default_rng(5).standard_normal((400, 200))
    A[17, 123] = np.nan
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    D.svd(400, 200, D.P(_t(A)), 200, 1, info)
    assert int(info[0].item()) == 2


@pytest.mark.parametrize("m,n,q_col_major,zero_cols", [(2000, 300, False, False),
                                                       (300, 2000, False, False),
                                                       (1000, 1000, False, False),
                                                       (700, 171, True, False),
                                                       (1500, 260, False, True)])
def test_qr_blocked_with_q(m, n, q_col_major, zero_cols):
    """Q and R of large matrices on the blocked path (Q accumulated from the
    kept panels in reverse, as dorgqr): Q R = A, Q^T Q = I, R upper triangular,
    |diag R| equal to LAPACK's; Q also written column-major (strided output)."""
    print(f"qr blocked with Q {m}x{n} col_major={q_col_major} zero_cols={zero_cols}", flush=True)
    D = _dev()
    rng = np.random.default_rng(7 * m + n)
    A = rng.standard_normal((m, n))
    if zero_cols:
        A[:, 50:80] = 0.0
    k = min(m, n)
    Q = D.empty(m, k)
    R = D.empty(k, n)
    if q_col_major:
        D.qr(m, n, D.P(_t(A)), n, 1, D.P(Q), 1, m, D.P(R), n, 1)
        Qh = _h(Q).reshape(-1)[:m * k].reshape(k, m).T
    else:
        D.qr(m, n, D.P(_t(A)), n, 1, D.P(Q), k, 1, D.P(R), n, 1)
        Qh = _h(Q)
    Rh = _h(R)
    np.testing.assert_allclose(Qh @ Rh, A, atol=1e-13 * np.abs(A).max() * max(m, n))
    np.testing.assert_allclose(Qh.T @ Qh, np.eye(k), atol=1e-13 * max(m, n))
    assert np.allclose(np.tril(Rh, -1), 0.0)
    ref = np.linalg.qr(A, mode="r")
    np.testing.assert_allclose(np.abs(np.diag(Rh)), np.abs(np.diag(ref)), rtol=1e-10,
                               atol=1e-12 * np.abs(np.diag(ref)).max())


@pytest.mark.parametrize("m,n", [(600, 200), (200, 600)])
def test_large_svd_and_qr_on_transposed_views(m, n):
    """The multi-workgroup SVD and the blocked QR read A through strides: a
    transposed view (column-major storage) gives the same factors as the
    row-major copy (the pre-sort and the QR see the same matrix)."""
    print(f"large svd/qr transposed view {m}x{n}", flush=True)
    D = _dev()
    rng = np.random.default_rng(m * 3 + n)
    A = rng.standard_normal((m, n)) * np.logspace(0, -6, n)[None, :]
    At = _t(np.ascontiguousarray(A.T))  # n x m row-major = A column-major
    Ar = _t(A)
    k = min(m, n)
    info = torch.zeros(2, dtype=torch.int32, device=D.device())
    U1, S1, V1 = D.svd(m, n, D.P(Ar), n, 1, info)
    assert int(info[0].item()) == 0
    U2, S2, V2 = D.svd(m, n, D.P(At), 1, m, info)
    assert int(info[0].item()) == 0
    np.testing.assert_array_equal(_h(S1), _h(S2))
    np.testing.assert_array_equal(_h(U1), _h(U2))
    np.testing.assert_array_equal(_h(V1), _h(V2))
    np.testing.assert_allclose(_h(S1), np.linalg.svd(A, compute_uv=False), rtol=1e-10,
                               atol=1e-13 * np.linalg.norm(A, 2))
    Q1, R1 = D.empty(m, k), D.empty(k, n)
    Q2, R2 = D.empty(m, k), D.empty(k, n)
    D.qr(m, n, D.P(Ar), n, 1, D.P(Q1), k, 1, D.P(R1), n, 1)
    D.qr(m, n, D.P(At), 1, m, D.P(Q2), k, 1, D.P(R2), n, 1)
    np.testing.assert_array_equal(_h(R1), _h(R2))
    np.testing.assert_array_equal(_h(Q1), _h(Q2))
    np.testing.assert_allclose(_h(Q1) @ _h(R1), A, atol=1e-13 * np.abs(A).max() * max(m, n))
