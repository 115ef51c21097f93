"""QTT conversions on the device path: ``tt_to_qtt`` / ``qtt_to_tt`` against the
reference's fixtures (tests/golden/qtt_*.npz, written by the reference) and the
contract of the reference's own tests (tests/test_tensor_train.py:392-620,
824-857: round trips, error bound, canonical form, bond floor, mode sizes that
need no split, regrouping, groups of 33+ cores, exact ranks, validation)."""

import numpy as np
import pytest

from conftest import load_golden, unpack_cores

pytestmark = pytest.mark.gpu

QTT_CASES = {
    "qtt_tt8": (dict(error=0.2), [8] * 4, dict(error=0.05)),
    "qtt_ttm4": (dict(error=1e-3, max_rank=4), [(4, 4)] * 3, dict(error=1e-3, max_rank=3)),
    "qtt_mode3": (dict(mode_size=3, error=1e-2), [3, 3, 27, 3], dict(error=1e-2)),
    "qtt_regroup": (None, [4, 16, 2, 2], dict(error=1e-2)),
}


def _tx():
    import ttnumpy_amd as tx
    return tx


def _rel(a, b):
    a, b = np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", sorted(QTT_CASES))
def test_qtt_conversions_match_reference(name):
    print(f"qtt reference case {name}", flush=True)
    tx = _tx()
    to_kw, idx, back_kw = QTT_CASES[name]
    g = load_golden(name)
    t = tx.TensorTrain(unpack_cores(g, "in_shapes", "in_flat"))
    q = t.tt_to_qtt(**to_kw) if to_kw is not None else t
    back = q.qtt_to_tt(idx, **back_kw)
    assert q.bonds == list(g["qtt_bonds"])
    assert _rel(q.get_full_tensor(), g["qtt_full"]) < 1e-10
    assert back.bonds == list(g["back_bonds"])
    assert back.orth_center == int(g["back_orth_center"])
    assert back.physical_dims() == [tuple(x) if isinstance(x, tuple) else x for x in idx]
    assert _rel(back.get_full_tensor(), g["back_full"]) < 1e-10


def test_qtt_round_trips():
    print("qtt round trips", flush=True)
    tx = _tx()
    for modes, ms, seed in (([8] * 4, 2, 115), ([27] * 4, 3, 116)):
        t = tx.rand(modes, ranks=5, seed=seed)
        q = t.tt_to_qtt(mode_size=ms, error=1e-10)
        assert q.physical_dims() == [ms] * 12
        assert _rel(q.get_full_tensor(), t.get_full_tensor()) < 1e-10
        back = q.qtt_to_tt(modes, error=1e-10)
        assert back.physical_dims() == modes
        assert _rel(back.get_full_tensor(), t.get_full_tensor()) < 1e-10
    t = tx.rand([(8, 8)] * 3, ranks=5, seed=117)
    q = t.tt_to_qtt(error=1e-10)
    assert q.physical_dims() == [(2, 2)] * 9
    assert _rel(q.get_full_tensor(), t.get_full_tensor()) < 1e-10
    assert _rel(q.qtt_to_tt([(8, 8)] * 3, error=1e-10).get_full_tensor(), t.get_full_tensor()) < 1e-10
    t = tx.rand([8] * 4, ranks=8, seed=118)
    x0 = t.get_full_tensor()
    for _ in range(5):
        t = t.tt_to_qtt(error=1e-10).qtt_to_tt([8] * 4, error=1e-10)
    assert _rel(t.get_full_tensor(), x0) < 1e-9


def test_qtt_error_bound_canonical_form_floor_and_source():
    print("qtt error bound and canonical form", flush=True)
    tx = _tx()
    from oracle import tt_oracle as O
    a = np.random.default_rng(0).standard_normal((8, 8, 8))
    # exact TT of the dense tensor by successive SVDs (input only)
    cores, r, rest = [], 1, a.reshape(1, -1)
    for n in (8, 8):
        u, s, vt = np.linalg.svd(rest.reshape(r * n, -1), full_matrices=False)
        cores.append(u.reshape(r, n, 1, -1))
        rest, r = s[:, None] * vt, u.shape[1]
    cores.append(rest.reshape(r, 8, 1, 1))
    t = tx.TensorTrain(cores)
    for err in (3e-1, 1e-1, 1e-2, 1e-4):
        q = t.tt_to_qtt(error=err)
        assert np.linalg.norm(q.get_full_tensor().reshape(-1) - a.reshape(-1)) <= err * np.linalg.norm(a)
        qo = O.tt_to_qtt(O.Train(cores), error=err)
        assert q.bonds == qo.bonds
    t = tx.rand([8, 8, 8], ranks=5, seed=119)
    before = [c.copy() for c in t.to_numpy()]
    q = t.tt_to_qtt(error=1e-10)
    assert t.orth_center is None
    assert all(np.array_equal(x, y) for x, y in zip(before, t.to_numpy()))
    assert q.orth_center == len(q) - 1
    for c in q.to_numpy()[:-1]:
        m = c.reshape(-1, c.shape[-1])
        np.testing.assert_allclose(m.T @ m, np.eye(m.shape[1]), atol=1e-10)
    full = np.linalg.norm(t.get_full_tensor())
    assert abs(q.tt_norm() - full) < 1e-10 * full
    q = tx.rand([16, 16], ranks=4, seed=120).tt_to_qtt(error=1e2)
    assert q.bonds == [1] * (len(q) + 1)


def test_qtt_mode_sizes_and_validation():
    print("qtt mode sizes and validation", flush=True)
    tx = _tx()
    t = tx.rand([2, 8], ranks=2, seed=121)
    q = t.tt_to_qtt(error=1e-10)
    assert q.physical_dims() == [2, 2, 2, 2]
    assert _rel(q.get_full_tensor(), t.get_full_tensor()) < 1e-10
    t = tx.rand([243, 3], ranks=2, seed=122)
    q = t.tt_to_qtt(mode_size=3, error=1e-10)
    assert q.physical_dims() == [3] * 6
    assert _rel(q.get_full_tensor(), t.get_full_tensor()) < 1e-10
    t = tx.rand([4, 4], ranks=2, seed=123)
    with pytest.raises(ValueError):
        t.tt_to_qtt(max_rank=0)
    with pytest.raises(ValueError):
        t.tt_to_qtt(mode_size=1)
    with pytest.raises(tx.DimensionMismatch):
        t.tt_to_qtt(mode_size=3)
    with pytest.raises(tx.DimensionMismatch):
        tx.rand([(4, 2), (4, 2)], ranks=2, seed=124).tt_to_qtt()
    q = tx.rand([4, 4, 4], ranks=3, seed=125).tt_to_qtt(error=1e-10)
    for idx in ([4, 4], [8, 8, 8], [3, 8, 8]):
        with pytest.raises(tx.DimensionMismatch):
            q.qtt_to_tt(idx, error=1e-10)
    ttm = tx.rand([(4, 4), (4, 4)], ranks=3, seed=126)
    qm = ttm.tt_to_qtt(error=1e-10)
    assert _rel(qm.qtt_to_tt([(4, 4), (4, 4)], error=1e-10).get_full_tensor(),
                ttm.get_full_tensor()) < 1e-10
    with pytest.raises(tx.DimensionMismatch):
        qm.qtt_to_tt([(4, 2), (4, 8)], error=1e-10)


def test_qtt_regrouping_and_long_groups():
    print("qtt regrouping and long groups", flush=True)
    tx = _tx()
    t = tx.rand([4, 4, 4], ranks=4, seed=0)
    q = t.tt_to_qtt(error=1e-10)
    for idx in ([8, 8], [2, 8, 4], [2] * 6):
        back = q.qtt_to_tt(idx, error=1e-10)
        assert back.physical_dims() == idx
        assert _rel(back.get_full_tensor(), t.get_full_tensor()) < 1e-10
    for qtt, idx in (
        (tx.rand([2] * 10 + [1] * 30 + [2] * 3, ranks=3, seed=7), [1024, 8]),
        (tx.rand([(2, 2)] * 4 + [(1, 1)] * 30 + [(2, 1)] * 3 + [(2, 2)] * 2, ranks=2, seed=9),
         [(128, 16), (4, 4)]),
    ):
        back = qtt.qtt_to_tt(idx)
        assert back.physical_dims() == idx
        ref = qtt.matricize()
        np.testing.assert_allclose(back.matricize(), ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_qtt_exact_ranks_at_default_error():
    print("qtt exact ranks", flush=True)
    tx = _tx()
    size = 256
    g = np.arange(size) / size
    t = tx.TensorTrain([np.stack([np.sin(10 * g), np.cos(10 * g)], 1).reshape(1, size, 1, 2),
                        np.stack([np.cos(10 * g), np.sin(10 * g)]).reshape(2, size, 1, 1)])
    q = t.tt_to_qtt()
    assert max(q.bonds) == 2
    assert _rel(q.get_full_tensor(), t.get_full_tensor()) < 1e-10
    back = q.qtt_to_tt([size, size])
    assert back.bonds == t.bonds
    assert _rel(back.get_full_tensor(), t.get_full_tensor()) < 1e-10
    g = np.arange(64) / 64
    s = tx.TensorTrain([np.stack([np.sin(10 * g), np.cos(10 * g)], 1).reshape(1, 64, 1, 2),
                        np.stack([np.cos(10 * g), np.sin(10 * g)]).reshape(2, 64, 1, 1)])
    assert max(s.tt_to_qtt(error=0.0).bonds) == 2
