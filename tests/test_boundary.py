"""The C-ABI boundary: the library loads and exports every symbol ``include/*.h`` declares.

CPU only: nothing here launches a kernel.
"""

import os
import re

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def _declared():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                   if f.endswith(".h"))
    return sorted(set(re.findall(r"^\s*(?:int|int64_t)\s+(ttx_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "ttx_gemm_f64" in names and "ttx_svd_f64" in names and len(names) > 15


def test_library_exports_every_declared_symbol():
    from ttnumpy_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libttx.so not built (run __graft_entry__.build())")
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) == set(_lib.exported_symbols())


def test_size_queries_are_host_only():
    from ttnumpy_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libttx.so not built")
    lib = _lib.load()
    assert lib.ttx_version() == 100
    assert lib.ttx_qr_work_size(10, 4) == 10 * 4 + 4 + 10 * 4
    assert lib.ttx_svd_work_size(3, 5) >= 2 * 15 + 2 * 9


def test_no_cpu_fallback_without_device():
    """Without a HIP device the package refuses to compute instead of falling back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from ttnumpy_amd import device as D
    with pytest.raises(RuntimeError):
        D.device()


def test_settings_validation_messages():
    from ttnumpy_amd.solvers.common import _SolverSettings
    kw = dict(max_sweeps=3, residual_tol=None, stagnation_factor="auto", local_solver="auto",
              symmetric=True, direct_solve_size=1000, return_info=False)
    with pytest.raises(ValueError, match="residual_tol must be positive"):
        _SolverSettings.resolve(**{**kw, "residual_tol": -1.0})
    with pytest.raises(ValueError, match="local_solver must be one of"):
        _SolverSettings.resolve(**{**kw, "local_solver": "cg"})
    with pytest.raises(ValueError, match="stagnation_factor"):
        _SolverSettings.resolve(**{**kw, "stagnation_factor": 1.5})
    s = _SolverSettings.resolve(**{**kw, "residual_tol": 1e-8}, truncation_steps=30)
    assert s.local_rtol == pytest.approx(1e-9)
    assert s.stagnation_factor == 0.9
    assert s.truncation_tol == pytest.approx(1e-8 / 30 ** 0.5)
    assert s.use_iterative(1001) and not s.use_iterative(1000)
    assert str(s.iterative_solver) == "minres"


def test_solver_enums_stringify_to_their_value():
    """str-enums: str() and format() give the bare value (the reference's
    SolverInfo.stop_reason reaches reports through str())."""
    from ttnumpy_amd.solvers.common import LocalSolver, StopReason, TruncationMode
    for enum_type in (StopReason, LocalSolver, TruncationMode):
        for member in enum_type:
            assert str(member) == member.value
            assert f"{member}" == member.value
