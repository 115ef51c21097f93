"""TensorTrain orthonormalization, SVD compression and truncation on the
device path: the contract of the reference's own tests
(tests/test_tensor_train.py: orthonormalize forms and centres, copy
semantics, partial sweeps, centre validation, compress_svd values and
centres, truncate without inplace)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tx():
    import ttnumpy_amd as tx
    return tx


def _left_ok(core, tol=1e-12):
    m = core.reshape(-1, core.shape[-1])
    np.testing.assert_allclose(m.T @ m, np.eye(m.shape[1]), atol=tol)


def _right_ok(core, tol=1e-12):
    m = core.reshape(core.shape[0], -1)
    np.testing.assert_allclose(m @ m.T, np.eye(m.shape[0]), atol=tol)


def test_orthonormalize_forms_and_centres():
    print("tt orthonormalize forms", flush=True)
    tx = _tx()
    t = tx.rand(((2, 3), (2, 3), (2, 3), (2, 3)), ranks=2, seed=3)
    full = t.get_full_tensor()
    d = t.n_cores
    left = t.orthonormalize("left")
    assert left.orth_center == d - 1
    for c in left.to_numpy()[:-1]:
        _left_ok(c)
    right = t.orthonormalize("right")
    assert right.orth_center == 0
    for c in right.to_numpy()[1:]:
        _right_ok(c)
    mix = t.orthonormalize("mix")
    assert mix.orth_center == d // 2
    cores = mix.to_numpy()
    for i in range(mix.orth_center):
        _left_ok(cores[i])
    for i in range(mix.orth_center + 1, d):
        _right_ok(cores[i])
    for u in (left, right, mix):
        np.testing.assert_allclose(u.get_full_tensor(), full, rtol=1e-12, atol=1e-12 * np.abs(full).max())


def test_orthonormalize_copy_semantics():
    tx = _tx()
    t = tx.rand((2, 3, 4), ranks=5, seed=101)
    before = [c.copy() for c in t.to_numpy()]
    o = t.orthonormalize("left")
    assert o is not t and t.orth_center is None
    for c0, c1 in zip(before, t.to_numpy()):
        np.testing.assert_array_equal(c0, c1)
    again = o.orthonormalize("left")  # already satisfied: still an independent copy
    assert again is not o and again.cores[0] is not o.cores[0]
    assert again.orth_center == o.orth_center


def test_orthonormalize_partial_sweeps_and_centre_checks():
    print("tt orthonormalize partial", flush=True)
    tx = _tx()
    t = tx.rand((2, 3, 4, 5), ranks=6, seed=102)
    full = t.get_full_tensor()
    part = t.orthonormalize("left", center_position=1)
    assert part.orth_center is None
    for u in (part, t.orthonormalize("right", center_position=2),
              t.orthonormalize("mix", center_position=1)):
        assert u.bonds == [c.shape[0] for c in u.to_numpy()] + [1]
        np.testing.assert_allclose(u.get_full_tensor(), full, rtol=1e-11, atol=1e-12 * np.abs(full).max())
    moved = t.orthonormalize("mix", center_position=1).orthonormalize("left", center_position=2)
    assert moved.orth_center == 2
    assert moved.tt_norm() == pytest.approx(np.linalg.norm(full), rel=1e-12)
    t3 = tx.rand((2, 3, 4), ranks=5, seed=103)
    for canonical in ("left", "right", "mix"):
        for bad in (3, -1):
            with pytest.raises(ValueError):
                t3.orthonormalize(canonical, center_position=bad)


def test_compress_svd_values_bounds_and_centres():
    print("tt compress_svd", flush=True)
    tx = _tx()
    t = tx.rand((2, 3, 4), ranks=5, seed=104)
    s = t.orthonormalize("left").compress_svd(1e-10)
    np.testing.assert_allclose(s.get_full_tensor(), t.get_full_tensor(), rtol=1e-9)
    assert s.orth_center == 0
    m = tx.rand(((2, 2), (3, 3), (4, 4)), ranks=6, seed=105)
    s2 = m.orthonormalize("right").compress_svd(1e-10, 3)
    assert max(s2.bonds) < 4 and s2.orth_center == 2


def test_truncate_copy_and_left_leaning_centre():
    print("tt truncate copy", flush=True)
    tx = _tx()
    t = tx.rand((3, 4, 5, 4, 3), ranks=4, seed=106)
    before = [c.copy() for c in t.to_numpy()]
    full = t.get_full_tensor()
    out = t.truncate(error=1e-12)
    assert out is not t
    for c0, c1 in zip(before, t.to_numpy()):
        np.testing.assert_array_equal(c0, c1)
    np.testing.assert_allclose(out.get_full_tensor(), full, rtol=1e-10, atol=1e-12 * np.abs(full).max())
    leaning = t.orthonormalize("left", center_position=1)
    out2 = leaning.truncate(error=1e-12)
    np.testing.assert_allclose(out2.get_full_tensor(), full, rtol=1e-10, atol=1e-12 * np.abs(full).max())


@pytest.mark.parametrize("max_rank", [None, 3])
def test_truncate_zero_train_keeps_orthonormal_cores(max_rank):
    """A train with a zero core: every SVD of the rounding has only zero
    singular values and keeps one (truncated_bond's floor, utils.py:22-69);
    its singular vectors are orthonormal, as LAPACK's are, so every core but
    the centre stays orthonormal, with the oracle's bonds and a zero tensor."""
    print(f"tt truncate zero train max_rank={max_rank}", flush=True)
    from oracle import tt_oracle as O
    tx = _tx()
    rng = np.random.default_rng(23)
    modes, ranks = [3, 4, 2, 5, 3], [1, 3, 4, 4, 2, 1]
    cores = [rng.standard_normal((ranks[i], modes[i], 1, ranks[i + 1])) for i in range(5)]
    cores[2][:] = 0.0
    out = tx.TensorTrain(cores).truncate(error=0.0, max_rank=max_rank)
    ref = O.truncate(O.Train(cores), 0.0, max_rank)
    assert out.bonds == ref.bonds == [1] * 6
    assert out.orth_center == 0
    got = out.to_numpy()
    for k in range(1, 5):
        _right_ok(got[k])
    assert np.all(out.get_full_tensor() == 0.0)


def test_add_with_randomized_truncation_passes_seed_and_options():
    """add(truncate=True, method="randomized", seed, options) rounds like
    truncate(method="randomized") with the same seed and options."""
    print("tt add randomized", flush=True)
    tx = _tx()
    a = tx.rand((3, 4, 2, 5), ranks=4, seed=21)
    b = tx.rand((3, 4, 2, 5), ranks=4, seed=22)
    opts = {"oversampling": 1}
    x = a.add(b, truncate=True, max_rank=3, method="randomized", seed=5, options=opts)
    y = a.add(b).truncate(max_rank=3, method="randomized", seed=5, options=opts)
    assert x.bonds == y.bonds
    np.testing.assert_allclose(x.get_full_tensor(), y.get_full_tensor(), rtol=1e-12, atol=1e-14)
    with pytest.raises(ValueError, match="needs max_rank"):
        a.add(b, truncate=True, method="randomized")


def test_elementwise_product_and_its_rounding():
    """x * y: bonds multiply, entries multiply (against the dense product);
    multiply(truncate=True) rounds it like the oracle's truncate."""
    print("tt elementwise product", flush=True)
    import os, sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from oracle import tt_oracle as O
    from ttnumpy_amd import problems as P
    tx = _tx()
    ca = P.random_train_cores([3, 4, 2, 5, 3], 3, 31, normal=True)
    cb = P.random_train_cores([3, 4, 2, 5, 3], 4, 32, normal=True)
    x, y = tx.TensorTrain(ca), tx.TensorTrain(cb)
    xy = x * y
    assert xy.bonds == [p * q for p, q in zip(x.bonds, y.bonds)]
    np.testing.assert_allclose(xy.get_full_tensor(), x.get_full_tensor() * y.get_full_tensor(),
                               rtol=1e-12, atol=1e-12)
    r = x.multiply(y, truncate=True, error=1e-10)
    ro = O.truncate(O.tt_multiply(O.Train(ca), O.Train(cb)), 1e-10)
    assert r.bonds == ro.bonds
    np.testing.assert_allclose(r.get_full_tensor().reshape(-1), ro.full(), rtol=1e-9,
                               atol=1e-11 * np.abs(ro.full()).max())
    A = tx.TensorTrain(P.poisson_2d_cores(2))
    with pytest.raises(tx.DimensionMismatch):
        x.multiply(A)


def test_randomized_elementwise_product_matches_reference():
    """multiply(truncate=True, method="randomized") on the device: the
    reference's bonds, centre and full tensor (tests/golden/multiply_rand.npz)."""
    print("tt randomized elementwise product", flush=True)
    import os, sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from conftest import load_golden
    from ttnumpy_amd import problems as P
    tx = _tx()
    g = load_golden("multiply_rand")
    x = tx.TensorTrain(P.random_train_cores([3, 4, 2, 5, 3], 3, 31, normal=True))
    y = tx.TensorTrain(P.random_train_cores([3, 4, 2, 5, 3], 4, 32, normal=True))
    out = x.multiply(y, truncate=True, max_rank=4, method="randomized", seed=3,
                     options={"oversampling": 2})
    assert out.bonds == list(g["bonds"])
    assert out.orth_center == int(g["orth_center"])
    full = out.get_full_tensor().reshape(-1)
    np.testing.assert_allclose(full, g["full"], rtol=1e-9, atol=1e-11 * np.abs(g["full"]).max())
