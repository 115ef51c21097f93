"""ttx_svd_f64_noalloc (the SVD entry point that allocates nothing) against ttx_svd_f64.

Both are called through the C ABI on the same inputs.  The reference
ttx_svd_f64 gets a work buffer filled with NaN, so a read of work it did not
write first would show up; the no-alloc variant runs on its file-scope
scratch array, which keeps the previous call's contents.  The calls go from
larger to smaller shapes (multi-workgroup Jacobi, CholeskyQR2 path,
one-workgroup kernel, wide input), so every call after the first sees a
scratch array dirtied by a larger problem.  Results must be bit-identical.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _svd_pair(lib, D, A):
    m, n = A.shape
    k = min(m, n)
    dA = D.to_device(A)
    out = {}
    for name in ("ref", "noalloc"):
        U = torch.full((m, k), float("nan"), dtype=torch.float64, device=D.device())
        S = torch.full((k,), float("nan"), dtype=torch.float64, device=D.device())
        Vt = torch.full((k, n), float("nan"), dtype=torch.float64, device=D.device())
        info = torch.full((1,), -1, dtype=torch.int32, device=D.device())
        if name == "ref":
            work = torch.full((int(lib.ttx_svd_work_size(m, n)),), float("nan"),
                              dtype=torch.float64, device=D.device())
            rc = lib.ttx_svd_f64(D.stream(), m, n, D.P(dA), n, 1, D.P(U), k, 1, D.P(S), D.P(Vt),
                                 n, 1, D.P(work), info.data_ptr())
        else:
            rc = lib.ttx_svd_f64_noalloc(D.stream(), m, n, D.P(dA), n, 1, D.P(U), k, 1, D.P(S),
                                         D.P(Vt), n, 1, info.data_ptr())
        assert rc == 0, (name, rc)
        torch.cuda.synchronize()
        out[name] = (U.cpu().numpy(), S.cpu().numpy(), Vt.cpu().numpy(), int(info.cpu()[0]))
    return out


def test_svd_noalloc_equals_svd_bitwise():
    from ttnumpy_amd import device as D
    lib = D.lib()
    assert lib.ttx_svd_noalloc_max_work() == 4 << 20
    rng = np.random.default_rng(7)
    # decreasing sizes: Q >= 139 (multi-workgroup Jacobi), the CholeskyQR2 path
    # (C2's 384 x 128), the one-workgroup kernel, a wide matrix
    for m, n in ((320, 160), (384, 128), (72, 40), (12, 30)):
        print(f"svd noalloc {m}x{n}", flush=True)
        A = rng.standard_normal((m, n)) * np.logspace(0, -4, n)
        got = _svd_pair(lib, D, A)
        (Ur, Sr, Vr, ir), (Un, Sn, Vn, inn) = got["ref"], got["noalloc"]
        assert ir == 0 and inn == 0
        assert np.array_equal(Sr, Sn) and np.array_equal(Ur, Un) and np.array_equal(Vr, Vn)
        # and it is an SVD of A
        np.testing.assert_allclose(Sn, np.linalg.svd(A, compute_uv=False), rtol=1e-11,
                                   atol=1e-13 * Sn[0])
        assert np.linalg.norm((Un * Sn) @ Vn - A) <= 1e-12 * np.linalg.norm(A)


def test_svd_noalloc_rejects_oversized_shape():
    from ttnumpy_amd import device as D
    lib = D.lib()
    m = n = 2000
    assert lib.ttx_svd_work_size(m, n) > lib.ttx_svd_noalloc_max_work()
    # rejected before any pointer is used or any kernel is launched
    rc = lib.ttx_svd_f64_noalloc(D.stream(), m, n, None, n, 1, None, n, 1, None, None, n, 1, None)
    assert rc == 7  # TTX_ESCRATCH
    with pytest.raises(ValueError, match="scratch"):
        from ttnumpy_amd._lib import check
        check(rc, "ttx_svd_f64_noalloc")
