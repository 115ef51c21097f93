"""Pins of the CPU oracle: reference-written golden vectors, known answers, brute force.

Everything here runs on the CPU.  The golden ``.npz`` files were written by
the reference library itself (``tests/golden/make_golden.py``); the known
answers are the values the reference's own tests assert
(``tests/test_utils.py:57-130``, ``tests/test_tensor_train.py:1789-1811``).
"""

import numpy as np
import pytest
from threadpoolctl import threadpool_limits

from conftest import load_golden, unpack_cores
from oracle import tt_oracle as O
from ttnumpy_amd import problems as P


# ---------------------------------------------------------------- truncated_bond
SV = np.array([4.0, 2.0, 1.0, 0.5])


@pytest.mark.parametrize(
    "values,delta,max_rank,expected",
    [
        (SV, 1.0, None, 3),
        (np.concatenate([[1.0], np.full(100, 0.099)]), 0.1, None, 100),
        (np.array([1.0, 0.05]), 0.1, None, 1),
        (np.array([1.0, 0.5]), 0.1, None, 2),
        (SV, 100.0, None, 1),
        (SV, 0.0, 2, 2),
        (np.array([1.0, 1e-17, 3e-18]), 0.0, None, 1),
        (np.array([1.0, 0.0, 0.0]), 0.0, None, 1),
        (np.array([1.0, 5e-17, 0.0]), 0.0, None, 1),
        (np.array([1.0, 1e-20, 1e-160]), 0.0, None, 1),
        (np.array([1e8, 1e-9]), 0.0, None, 1),
        (np.array([1e8, 1e-5]), 0.0, None, 2),
        (SV, 0.0, None, 4),
        (np.array([1.0, 1e-9]), 0.0, None, 2),
        (np.array([1.0, 1e-9], dtype=np.float32), 0.0, None, 1),
    ],
)
def test_truncated_bond_known_answers(values, delta, max_rank, expected):
    assert O.truncated_bond(values, delta, max_rank) == expected


def test_truncated_bond_rejects_bad_max_rank():
    for mr in (0, -3):
        with pytest.raises(ValueError):
            O.truncated_bond(SV, 1.0, mr)


# ---------------------------------------------------------------- truncate
def test_truncate_two_core_closed_form():
    rng = np.random.default_rng(19)
    left, _ = np.linalg.qr(rng.standard_normal((30, 6)))
    right, _ = np.linalg.qr(rng.standard_normal((30, 6)))
    sigma, mix = 0.5 ** np.arange(6), rng.standard_normal((6, 6))
    t = O.Train([((left * sigma) @ mix).reshape(1, 30, 1, 6),
                 np.linalg.solve(mix, right.T).reshape(6, 30, 1, 1)])
    full = t.full()
    best = np.linalg.norm(sigma[2:]) / np.linalg.norm(sigma)
    for canonical in ("auto", "left", "right"):
        capped = O.truncate(t, max_rank=2, canonical=canonical)
        err = np.linalg.norm(capped.full() - full) / np.linalg.norm(full)
        assert err == pytest.approx(best, rel=1e-10)
        assert O.truncate(t, error=0.2, canonical=canonical).bonds == [1, 3, 1]


@pytest.mark.parametrize("name,modes,rank,err,mr", [
    ("round_small_a", [3, 4, 2, 5, 3, 2], 6, 1e-6, None),
    ("round_small_b", [2] * 8, 7, 0.0, 5),
    ("round_small_c", [4] * 5, 9, 1e-2, None),
])
def test_truncate_matches_reference_small(name, modes, rank, err, mr):
    g = load_golden(name)
    s = O.tt_add(O.Train(P.random_train_cores(modes, rank, 11, normal=True)),
                 O.Train(P.random_train_cores(modes, rank, 12, normal=True)))
    out = O.truncate(s, err, mr)
    assert out.bonds == list(g["bonds"])
    assert out.orth_center == int(g["orth_center"])
    np.testing.assert_allclose(out.full(), g["full"], rtol=1e-9, atol=1e-11 * np.abs(g["full"]).max())
    # right-orthonormal cores after the right-to-left SVD sweep
    for c in out.cores[1:]:
        m = c.reshape(c.shape[0], -1)
        np.testing.assert_allclose(m @ m.T, np.eye(c.shape[0]), atol=1e-12)


def test_truncate_c2_spectra_match_reference():
    g = load_golden("c2_round_d40")
    s = O.tt_add(O.Train(P.random_train_cores([4] * 40, 64, 1, normal=True)),
                 O.Train(P.random_train_cores([4] * 40, 64, 2, normal=True)))
    spectra = []
    out = O.truncate(s, 1e-10, 96, spectra=spectra)
    assert out.bonds == list(g["bonds"])
    lens = g["spectra_lengths"]
    ref = np.split(g["spectra"], np.cumsum(lens)[:-1])
    assert len(spectra) == len(ref)
    for a, r in zip(spectra, ref):
        np.testing.assert_allclose(a, r, rtol=1e-10)


# ---------------------------------------------------------------- environments / local system
def _tiny_system(seed=0):
    rng = np.random.default_rng(seed)
    d, n = 4, 3
    A = O.Train([rng.standard_normal((r0, n, n, r1)) for r0, r1 in zip([1, 2, 3, 2], [2, 3, 2, 1])])
    x = O.Train([rng.standard_normal((r0, n, 1, r1)) for r0, r1 in zip([1, 3, 4, 2], [3, 4, 2, 1])])
    y = O.Train([rng.standard_normal((r0, n, 1, r1)) for r0, r1 in zip([1, 2, 3, 2], [2, 3, 2, 1])])
    return A, x, y, d, n


def _interface(t: O.Train, k: int, side: str) -> np.ndarray:
    """Dense left interface (prod modes[:k], bond k) or right (bond k+1, prod modes[k+1:])."""
    if side == "left":
        m = np.ones((1, 1))
        for c in t.cores[:k]:
            m = (m @ c[:, :, 0, :].reshape(c.shape[0], -1)).reshape(-1, c.shape[3])
        return m
    m = np.ones((1, 1))
    for c in reversed(t.cores[k + 1:]):
        m = (c[:, :, 0, :].reshape(-1, c.shape[3]) @ m).reshape(c.shape[0], -1)
    return m


def test_envs_and_local_operator_brute_force():
    """Psi . A_k . Phi equals X_{!=k}^T A X_{!=k} from dense interface matrices."""
    A, x, y, d, n = _tiny_system()
    Ad = A.dense_matrix()
    for k in range(d):
        eL = [None] * d
        eR = [None] * d
        for i in range(k + 1):
            O.build_env(eL, i, "left", x, y, A)
        for i in range(d - 1, k - 1, -1):
            O.build_env(eR, i, "right", x, y, A)
        # frame x (rows), target y (cols): M = Xf^T A Xt restricted to site k
        XL, XR = _interface(x, k, "left"), _interface(x, k, "right")
        YL, YR = _interface(y, k, "left"), _interface(y, k, "right")
        nl, nr = XL.shape[0], XR.shape[1]
        frame = np.einsum("pa,bq->paq b".replace(" ", ""), XL, XR) if False else None
        # rows (f_l, i, f_r) -> full index (p, i, q)
        Pf = np.einsum("pa,ij,bq->aibpjq", XL, np.eye(n), XR).reshape(XL.shape[1] * n * XR.shape[0], -1)
        Pt = np.einsum("pa,ij,bq->aibpjq", YL, np.eye(n), YR).reshape(YL.shape[1] * n * YR.shape[0], -1)
        dense = Pf @ Ad @ Pt.T
        M = O.local_matrix(eL[k], A.cores[k], eR[k])
        np.testing.assert_allclose(M, dense, rtol=1e-11, atol=1e-11 * np.abs(dense).max())
        v = np.random.default_rng(k).standard_normal((M.shape[1], 3))
        np.testing.assert_allclose(O.local_apply(eL[k], A.cores[k], eR[k], v), M @ v,
                                   rtol=1e-11, atol=1e-11 * np.abs(M).max())
        # overlap envs and the local rhs: b_loc = Pf @ y_full
        bl = [None] * d
        br = [None] * d
        for i in range(k + 1):
            O.build_env(bl, i, "left", x, y)
        for i in range(d - 1, k - 1, -1):
            O.build_env(br, i, "right", x, y)
        rhs = O.local_rhs(bl[k], y.cores[k][:, :, 0, :], br[k])
        Py = np.einsum("pa,ij,bq->pjq", YL, np.eye(n), YR) if False else None
        yfull = y.full()
        np.testing.assert_allclose(rhs, Pf @ yfull, rtol=1e-11, atol=1e-11 * np.abs(yfull).max())


def test_relative_residual_dense():
    A, x, y, d, n = _tiny_system(3)
    Ad = A.dense_matrix()
    r = np.linalg.norm(Ad @ x.full() - y.full()) / np.linalg.norm(y.full())
    assert O.relative_residual(A, y, x) == pytest.approx(r, rel=1e-10)


# ---------------------------------------------------------------- solver trajectories
def _check_traj(x, info, g, rtol_res=1e-5):
    assert x.bonds == list(g["bonds"])
    assert info.num_sweeps_run == int(g["num_sweeps_run"])
    assert info.stop_reason == str(g["stop_reason"])
    assert info.max_bonds == list(g["max_bonds"])
    np.testing.assert_allclose(info.residuals, g["residuals"], rtol=rtol_res)
    ref = O.Train(unpack_cores(g))
    diff = O.tt_norm(O.tt_add(x, O.tt_scale(ref, -1.0))) / O.tt_norm(ref)
    assert diff < 1e-10


def test_amen_poisson64_matches_reference():
    g = load_golden("poisson64_amen")
    A = O.Train(P.poisson_2d_cores(6))
    b = O.Train(P.ones_cores([2] * 12))
    x0 = O.Train(P.random_train_cores([2] * 12, 2, 0))
    x, info = O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0)
    _check_traj(x, info, g)


def test_amen_poisson64_minres_matches_reference():
    g = load_golden("poisson64_amen_minres")
    A = O.Train(P.poisson_2d_cores(6))
    b = O.Train(P.ones_cores([2] * 12))
    x0 = O.Train(P.random_train_cores([2] * 12, 2, 0))
    x, info = O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0,
                     direct_solve_size=16)
    _check_traj(x, info, g)
    assert info.local_solves == int(g["local_solves"])


def test_als_lap1d_matches_reference():
    g = load_golden("lap1d6_als")
    h = 1.0 / (2 ** 6 + 1)
    A = O.Train(P.laplacian_1d_cores(6, h))
    b = O.Train(P.ones_cores([2] * 6))
    x0 = O.orthonormalize(O.Train(P.ones_cores([2] * 6, 4)), "right")
    x, info = O.als(A, b, x0, max_sweeps=4, residual_tol=1e-10)
    _check_traj(x, info, g, rtol_res=1e-3)
    # dense pin: the solve reaches the exact discrete solution
    xd = np.linalg.solve(A.dense_matrix(), b.full())
    assert np.linalg.norm(x.full() - xd) / np.linalg.norm(xd) < 1e-8


def _check_site_replay(states, A, b, settings):
    """Every hooked state, replayed by ``site_step``, gives exactly the next
    state's cores: the replay is the solver's own site step."""
    steps = 0
    for s, nxt in zip(states, states[1:]):
        out = O.site_step(s, A, b, settings)
        if "trace" not in out:
            assert s.direction == "backward" and s.k == 0
            continue
        nb = out["neighbour"]
        for got, want in ((out["x_k"], nxt.x[s.k]), (out["x_nb"], nxt.x[nb]),
                          (out["z_k"], nxt.z[s.k]), (out["z_nb"], nxt.z[nb])):
            assert got.shape == want.shape and np.array_equal(got, want), (s.sweep, s.direction, s.k)
        steps += 1
    return steps


def test_amen_c1_poisson256_matches_reference():
    g = load_golden("c1_poisson256_amen")
    A = O.Train(P.poisson_2d_cores(8))
    b = O.Train(P.ones_cores([2] * 16))
    x0 = O.Train(P.random_train_cores([2] * 16, 2, 0))
    states = []
    x, info = O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0,
                     hook=states.append)
    _check_traj(x, info, g)
    # the hook sees every site step of the golden run: 6 sweeps x (15 + 16)
    assert len(states) == int(g["num_sweeps_run"]) * 31
    st = O.resolve_settings(20, 1e-8, "auto", "auto", True, 1000, True, truncation_steps=30)
    assert _check_site_replay(states, A, b, st) == len(states) - int(g["num_sweeps_run"])


def test_site_replay_minres_path():
    """The replay of the iterative (MINRES) local solves, on the 64 x 64 golden run."""
    g = load_golden("poisson64_amen_minres")
    A = O.Train(P.poisson_2d_cores(6))
    b = O.Train(P.ones_cores([2] * 12))
    x0 = O.Train(P.random_train_cores([2] * 12, 2, 0))
    states = []
    x, info = O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0,
                     direct_solve_size=16, hook=states.append)
    _check_traj(x, info, g)
    st = O.resolve_settings(20, 1e-8, "auto", "auto", True, 16, True, truncation_steps=22)
    assert _check_site_replay(states, A, b, st) > 0


def test_poisson_operator_is_the_kronecker_sum():
    """The QTT builder reproduces the dense 5-point Laplacian (brute force, 8x8 grid)."""
    nb = 3
    A = O.Train(P.poisson_2d_cores(nb)).dense_matrix()
    m = 2 ** nb
    h = 1.0 / (m + 1)
    L = (2 * np.eye(m) - np.eye(m, k=1) - np.eye(m, k=-1)) / h ** 2
    dense = np.kron(L, np.eye(m)) + np.kron(np.eye(m), L)
    np.testing.assert_allclose(A, dense, rtol=1e-12, atol=1e-9)
    assert O.Train(P.poisson_2d_cores(nb)).bonds == [1, 4, 4, 2, 4, 4, 1]


RAND_CASES = [
    ("round_rand_a", [3, 4, 2, 5, 3, 2], 6, 3, 2, 0.0, None, 7),
    ("round_rand_b", [3, 4, 2, 5, 3, 2], 6, 4, 3, 1e-8, "right", 8),
    ("round_rand_c", [2] * 8, 7, 4, 10, 0.0, None, 9),
]


def _rand_input(modes, rank, pre):
    s = O.tt_add(O.Train(P.random_train_cores(modes, rank, 11, normal=True)),
                 O.Train(P.random_train_cores(modes, rank, 12, normal=True)))
    if pre is not None:
        O.orthonormalize(s, pre)
    return s


@pytest.mark.parametrize("name,modes,rank,mr,over,err,pre,seed", RAND_CASES)
def test_randomized_truncate_matches_reference(name, modes, rank, mr, over, err, pre, seed):
    """Sketch bonds below (a, b) and above (c) the train's ranks, both sweep
    directions (b starts right-orthonormal): the reference's bonds, centre
    and full tensor."""
    g = load_golden(name)
    out = O.truncate(_rand_input(modes, rank, pre), err, mr, method="randomized", seed=seed,
                     options={"oversampling": over})
    assert out.bonds == list(g["bonds"])
    assert out.orth_center == int(g["orth_center"])
    np.testing.assert_allclose(out.full(), g["full"], rtol=1e-9, atol=1e-11 * np.abs(g["full"]).max())


def test_randomized_sketch_is_exact_when_ranks_fit():
    """With every unfolding rank at most the sketch bond, the randomized sweep
    spans the train exactly, so rounding gives the direct result."""
    s = _rand_input([3, 4, 2, 5, 3], 3, None)   # sum of two rank-3 trains: ranks <= 6
    direct = O.truncate(s, 1e-12, 6)
    for seed in (0, 1):
        rnd = O.truncate(s, 1e-12, 6, method="randomized", seed=seed, options={"oversampling": 0})
        assert rnd.bonds == direct.bonds
        np.testing.assert_allclose(rnd.full(), direct.full(), rtol=1e-10,
                                   atol=1e-12 * np.abs(direct.full()).max())


def test_randomized_options_are_checked():
    s = _rand_input([2, 3, 2], 2, None)
    with pytest.raises(ValueError, match="needs max_rank"):
        O.truncate(s, method="randomized")
    with pytest.raises(ValueError, match="Unknown options"):
        O.truncate(s, max_rank=2, method="randomized", options={"over": 1})
    for bad in (-1, 1.5, "2"):
        with pytest.raises(ValueError, match="oversampling"):
            O.truncate(s, max_rank=2, method="randomized", options={"oversampling": bad})
    with pytest.raises(ValueError, match="Method"):
        O.truncate(s, max_rank=2, method="sketchy")


MALS_CASES = [
    ("mals_lap1d6", "lap1d", 6, dict(residual_tol=1e-10, max_sweeps=4, symmetric=True)),
    ("mals_poisson16", "poisson", 4, dict(residual_tol=1e-8, max_sweeps=6, symmetric=True)),
    ("mals_poisson16_minres", "poisson", 4,
     dict(residual_tol=1e-8, max_sweeps=6, symmetric=True, direct_solve_size=16)),
    ("mals_poisson8_threshold", "poisson", 3,
     dict(error_threshold=1e-9, max_sweeps=2, symmetric=True, stagnation_factor=None)),
]


def _mals_system(kind, n):
    if kind == "lap1d":
        return (O.Train(P.laplacian_1d_cores(n, 1.0 / (2 ** n + 1))), O.Train(P.ones_cores([2] * n)), n)
    return O.Train(P.poisson_2d_cores(n)), O.Train(P.ones_cores([2] * (2 * n))), 2 * n


@pytest.mark.parametrize("name,kind,n,kw", MALS_CASES)
def test_mals_matches_reference(name, kind, n, kw):
    """Two-site MALS: the reference's bonds, sweeps, stop reason and solution;
    residual histories to the accuracy the local solves leave (the last
    entries sit at round-off)."""
    g = load_golden(name)
    A, b, d = _mals_system(kind, n)
    x0 = O.Train(P.random_train_cores([2] * d, 2, 0))
    with threadpool_limits(1):
        x, info = O.mals(A, b, x0, **kw)
    assert x.bonds == list(g["bonds"])
    assert info.num_sweeps_run == int(g["num_sweeps_run"])
    assert info.stop_reason == str(g["stop_reason"])
    r, rg = np.array(info.residuals), g["residuals"]
    assert len(r) == len(rg)
    for a, c in zip(r, rg):
        if c > 1e-8:
            assert abs(a - c) <= 1e-5 * c
        elif c > 1e-12:  # set by the local solves' tolerance (1e-9 relative)
            assert abs(a - c) <= 1e-2 * c
        else:  # round-off level on both sides
            assert a <= 1e-12
    ref = g["full"]
    assert np.linalg.norm(x.full() - ref) <= 1e-8 * np.linalg.norm(ref)


def test_mals_two_site_operator_is_the_projection():
    """Brute force: the two-site local matrix of a random frame equals the
    dense operator projected onto span(left frame) x I x I x span(right frame)."""
    rng = np.random.default_rng(4)
    d, n = 4, 2
    A = O.Train([rng.standard_normal((r0, n, n, r1))
                 for r0, r1 in zip([1, 2, 3, 2], [2, 3, 2, 1])])
    x = O.Train([rng.standard_normal((r0, n, 1, r1))
                 for r0, r1 in zip([1, 2, 2, 2], [2, 2, 2, 1])])
    O.orthonormalize(x, "mix", 1)
    eL, eR = [None] * d, [None] * d
    for k in range(d - 1, -1, -1):
        O.build_env(eR, k, "right", x, x, A)
    for k in range(d):
        O.build_env(eL, k, "left", x, x, A)
    k = 1
    M = O.local_matrix(eL[k], O._merge_pair(A.cores[k], A.cores[k + 1]), eR[k + 1])
    # frames: left = core 0 (2 x 2 -> (n, r1)), right = core 3 ((r3, n))
    Lf = x.cores[0].reshape(n, -1)
    Rf = x.cores[3].reshape(-1, n)
    Pm = np.kron(np.kron(Lf, np.eye(n * n)), Rf.T)        # (16 x 2*4*2)
    dense = A.dense_matrix()
    np.testing.assert_allclose(M, Pm.T @ dense @ Pm, atol=1e-12)


def test_elementwise_product_closed_form():
    """full(a * b) = full(a) * full(b), for trains and for operator cores."""
    a = O.Train(P.random_train_cores([3, 4, 2, 3], 3, 1, normal=True))
    b = O.Train(P.random_train_cores([3, 4, 2, 3], 2, 2, normal=True))
    ab = O.tt_multiply(a, b)
    np.testing.assert_allclose(ab.full(), a.full() * b.full(), rtol=1e-12, atol=1e-13)
    assert ab.bonds == [x * y for x, y in zip(a.bonds, b.bonds)]
    A = O.Train(P.poisson_2d_cores(2))
    np.testing.assert_allclose(O.tt_multiply(A, A).dense_matrix(), A.dense_matrix() ** 2, atol=1e-12)


def test_randomized_elementwise_product_matches_reference():
    """multiply(truncate=True, method="randomized"): the reference's bonds,
    centre and full tensor; with a sketch covering the product's ranks it
    equals the rounding of the exact product."""
    g = load_golden("multiply_rand")
    a = O.Train(P.random_train_cores([3, 4, 2, 5, 3], 3, 31, normal=True))
    b = O.Train(P.random_train_cores([3, 4, 2, 5, 3], 4, 32, normal=True))
    out = O.multiply_rounded(a, b, max_rank=4, method="randomized", seed=3,
                             options={"oversampling": 2})
    assert out.bonds == list(g["bonds"])
    assert out.orth_center == int(g["orth_center"])
    np.testing.assert_allclose(out.full(), g["full"], rtol=1e-9, atol=1e-11 * np.abs(g["full"]).max())
    small_a = O.Train(P.random_train_cores([2, 3, 2], 2, 5, normal=True))
    small_b = O.Train(P.random_train_cores([2, 3, 2], 2, 6, normal=True))
    exact = O.multiply_rounded(small_a, small_b, error=1e-12, max_rank=4)
    rnd = O.multiply_rounded(small_a, small_b, error=1e-12, max_rank=4, method="randomized",
                             seed=0, options={"oversampling": 0})
    assert rnd.bonds == exact.bonds
    np.testing.assert_allclose(rnd.full(), exact.full(), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- QTT conversions
QTT_CASES = {
    # name: (tt_to_qtt kwargs or None, qtt_to_tt physical indexes, qtt_to_tt kwargs)
    "qtt_tt8": (dict(error=0.2), [8] * 4, dict(error=0.05)),
    "qtt_ttm4": (dict(error=1e-3, max_rank=4), [(4, 4)] * 3, dict(error=1e-3, max_rank=3)),
    "qtt_mode3": (dict(mode_size=3, error=1e-2), [3, 3, 27, 3], dict(error=1e-2)),
    "qtt_regroup": (None, [4, 16, 2, 2], dict(error=1e-2)),
}


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def _full(t):
    """The reference's get_full_tensor order: a TTM as its dense matrix."""
    return t.dense_matrix().reshape(-1) if any(c > 1 for c in t.col_modes) else t.full()


@pytest.mark.parametrize("name", sorted(QTT_CASES))
def test_qtt_conversions_match_reference(name):
    """tt_to_qtt then qtt_to_tt reproduce the reference's bonds and tensors
    (fixtures written by the reference, tests/golden/qtt_*.npz)."""
    to_kw, idx, back_kw = QTT_CASES[name]
    g = load_golden(name)
    t = O.Train(unpack_cores(g, "in_shapes", "in_flat"))
    with threadpool_limits(1):
        q = O.tt_to_qtt(t, **to_kw) if to_kw is not None else t
        back = O.qtt_to_tt(q, idx, **back_kw)
    assert q.bonds == list(g["qtt_bonds"])
    assert _rel(_full(q), g["qtt_full"]) < 1e-10
    assert back.bonds == list(g["back_bonds"])
    assert back.orth_center == int(g["back_orth_center"])
    assert _rel(_full(back), g["back_full"]) < 1e-10


def _dense_tt(tensor):
    """Exact TT of a small dense tensor by successive SVDs (test input only)."""
    dims = tensor.shape
    cores, r, rest = [], 1, tensor.reshape(1, -1)
    for n in dims[:-1]:
        u, s, vt = np.linalg.svd(rest.reshape(r * n, -1), full_matrices=False)
        k = int(np.sum(s > 1e-14 * s[0]))
        cores.append(u[:, :k].reshape(r, n, 1, k))
        rest, r = s[:k, None] * vt[:k], k
    cores.append(rest.reshape(r, dims[-1], 1, 1))
    return O.Train(cores)


def test_qtt_round_trips_and_modes():
    """Round trips at 1e-10 for mode 2, mode 3 and a TTM; five cycles keep the
    tensor (the reference's test_qtt, tests/test_tensor_train.py:392-471)."""
    for modes, ms, seed in (([8] * 4, 2, 115), ([27] * 4, 3, 116)):
        t = O.Train(P.random_train_cores(modes, 5, seed))
        q = O.tt_to_qtt(t, mode_size=ms, error=1e-10)
        assert q.row_modes == [ms] * (3 * 4)
        assert _rel(q.full(), t.full()) < 1e-10
        back = O.qtt_to_tt(q, modes, error=1e-10)
        assert back.row_modes == modes
        assert _rel(back.full(), t.full()) < 1e-10
    rng = np.random.default_rng(117)
    b = [1, 5, 5, 1]
    t = O.Train([rng.random((b[i], 8, 8, b[i + 1])) for i in range(3)])
    q = O.tt_to_qtt(t, error=1e-10)
    assert q.row_modes == [2] * 9 and q.col_modes == [2] * 9
    assert _rel(_full(q), _full(t)) < 1e-10
    assert _rel(_full(O.qtt_to_tt(q, [(8, 8)] * 3, error=1e-10)), _full(t)) < 1e-10
    t = O.Train(P.random_train_cores([8] * 4, 8, 118))
    x0 = t.full()
    for _ in range(5):
        t = O.qtt_to_tt(O.tt_to_qtt(t, error=1e-10), [8] * 4, error=1e-10)
    assert _rel(t.full(), x0) < 1e-9


def test_qtt_error_bound_canonical_form_and_floor():
    """||qtt - A|| <= error ||A||; left-orthonormal cores with the centre last
    and the norm in it; error 1e2 floors every bond at 1
    (tests/test_tensor_train.py:473-517)."""
    a = np.random.default_rng(0).standard_normal((8, 8, 8))
    t = _dense_tt(a)
    for err in (3e-1, 1e-1, 1e-2, 1e-4):
        q = O.tt_to_qtt(t, error=err)
        assert np.linalg.norm(q.full() - a.reshape(-1)) <= err * np.linalg.norm(a)
    t = O.Train(P.random_train_cores([8, 8, 8], 5, 119))
    q = O.tt_to_qtt(t, error=1e-10)
    assert q.orth_center == q.d - 1
    for c in q.cores[:-1]:
        m = c.reshape(-1, c.shape[-1])
        np.testing.assert_allclose(m.T @ m, np.eye(m.shape[1]), atol=1e-10)
    assert abs(O.tt_norm(q) - np.linalg.norm(t.full())) < 1e-10 * np.linalg.norm(t.full())
    before = [c.copy() for c in t.cores]
    O.tt_to_qtt(t, error=1e-10)
    assert all(np.array_equal(x, y) for x, y in zip(before, t.cores))
    q = O.tt_to_qtt(O.Train(P.random_train_cores([16, 16], 4, 120)), error=1e2)
    assert q.bonds == [1] * (q.d + 1)


def test_qtt_mixed_modes_and_validation():
    """Cores with no split, 243 = 3^5, and the argument checks
    (tests/test_tensor_train.py:519-552,593-620)."""
    t = O.Train(P.random_train_cores([2, 8], 2, 121))
    q = O.tt_to_qtt(t, error=1e-10)
    assert q.row_modes == [2, 2, 2, 2] and _rel(q.full(), t.full()) < 1e-10
    t = O.Train(P.random_train_cores([243, 3], 2, 122))
    q = O.tt_to_qtt(t, mode_size=3, error=1e-10)
    assert q.row_modes == [3] * 6 and _rel(q.full(), t.full()) < 1e-10
    t = O.Train(P.random_train_cores([4, 4], 2, 123))
    for kw in (dict(max_rank=0), dict(mode_size=1), dict(mode_size=3)):
        with pytest.raises(ValueError):
            O.tt_to_qtt(t, **kw)
    rng = np.random.default_rng(124)
    rect = O.Train([rng.random((1, 4, 2, 2)), rng.random((2, 4, 2, 1))])
    with pytest.raises(ValueError):
        O.tt_to_qtt(rect)
    q = O.tt_to_qtt(O.Train(P.random_train_cores([4, 4, 4], 3, 125)), error=1e-10)
    for idx in ([4, 4], [8, 8, 8], [3, 8, 8]):
        with pytest.raises(ValueError):
            O.qtt_to_tt(q, idx, error=1e-10)


def test_qtt_regrouping_and_long_groups():
    """Groups of different sizes, and groups of 33+ cores with modes of 1
    (tests/test_tensor_train.py:554-591)."""
    a = np.random.default_rng(0).standard_normal((4, 4, 4))
    q = O.tt_to_qtt(_dense_tt(a), error=1e-10)
    for idx in ([8, 8], [2, 8, 4], [2] * 6):
        back = O.qtt_to_tt(q, idx, error=1e-10)
        assert back.row_modes == idx
        assert _rel(back.full(), a.reshape(-1)) < 1e-10
    t = O.Train(P.random_train_cores([2] * 10 + [1] * 30 + [2] * 3, 3, 7))
    back = O.qtt_to_tt(t, [1024, 8])
    assert back.row_modes == [1024, 8]
    np.testing.assert_allclose(back.full(), t.full(), rtol=1e-12, atol=1e-14 * np.abs(t.full()).max())


def test_merge_cores_against_dense_matrix():
    """A merged TTM core is the dense matrix of the group: rows and columns in
    the cores' order (brute force with Kronecker products)."""
    rng = np.random.default_rng(5)
    mats = [rng.standard_normal((2, 3)), rng.standard_normal((4, 1)), rng.standard_normal((3, 2))]
    cores = [m.reshape(1, *m.shape, 1) for m in mats]
    merged = O.merge_cores(cores)
    np.testing.assert_allclose(merged[0, :, :, 0], np.kron(np.kron(mats[0], mats[1]), mats[2]),
                               rtol=1e-14)


def test_qtt_sine_of_sum_keeps_exact_ranks():
    """sin(10 (x + y)) on 256 x 256 has TT and QTT rank 2: kept at the default
    error, and at error 0 (tests/test_tensor_train.py:824-857)."""
    size = 256
    g = np.arange(size) / size
    t = O.Train([np.stack([np.sin(10 * g), np.cos(10 * g)], 1).reshape(1, size, 1, 2),
                 np.stack([np.cos(10 * g), np.sin(10 * g)]).reshape(2, size, 1, 1)])
    q = O.tt_to_qtt(t)
    assert max(q.bonds) == 2
    assert _rel(q.full(), t.full()) < 1e-10
    back = O.qtt_to_tt(q, [size, size])
    assert back.bonds == [1, 2, 1]
    assert _rel(back.full(), t.full()) < 1e-10
    assert max(O.tt_to_qtt(O.Train([c[:, ::4] for c in t.cores]), error=0.0).bonds) == 2


def test_poisson_exact_solution_closed_form():
    """The DST closed form used to pin AMEn solutions equals a dense solve of
    the QTT operator (16 x 16 grid), and the stencil equals that operator."""
    import poisson_exact as PE
    A = O.Train(P.poisson_2d_cores(4))
    M = A.dense_matrix()
    rng = np.random.default_rng(1)
    B = rng.standard_normal((16, 16))
    np.testing.assert_allclose(PE.solve(B).reshape(-1), np.linalg.solve(M, B.reshape(-1)),
                               rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(PE.apply(B).reshape(-1), M @ B.reshape(-1), rtol=1e-13, atol=1e-10)
    ev = np.linalg.eigvalsh(M)
    assert PE.lam_min(4) == pytest.approx(ev[0], rel=1e-12)


def test_c1_reference_solution_within_residual_bound():
    """The reference's own C1 solution (golden) is within ||Ax - b|| / lam_min
    of the exact discrete solution: the bound the device tests apply."""
    import poisson_exact as PE
    g = load_golden("c1_poisson256_amen")
    x = O.Train(unpack_cores(g)).full()
    err, bound, nrm = PE.error_bound(x, np.ones(2**16), 8)
    assert err <= bound * (1 + 1e-6)
    assert bound <= 1e-8 * np.sqrt(2**16) / PE.lam_min(8) * 1.01  # residual 8.2e-9 <= 1e-8
    assert err / nrm < 1e-7


@pytest.mark.parametrize("theta", [0.0, 0.37, 0.9])
def test_diffusion_sparse_operator_is_the_ttm(theta):
    """C4's QTT operator (16 x 16 grid) equals the sparse finite-volume matrix the
    device tests solve for the exact discrete solution, which is SPD with its
    smallest eigenvalue above the Laplacian's."""
    import poisson_exact as PE
    A = O.Train(P.diffusion_2d_cores(4, theta)).dense_matrix()
    S = PE.diffusion_sparse(4, theta).toarray()
    np.testing.assert_allclose(A, S, rtol=1e-12, atol=1e-9)
    assert np.linalg.eigvalsh(S)[0] >= PE.lam_min(4) * (1 - 1e-12)
