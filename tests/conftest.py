"""Shared pytest configuration: the ``gpu`` marker and fixture helpers."""

import os
import sys

# one BLAS thread for the oracle (see the note at the end of this file); set
# before NumPy/SciPy load their OpenBLAS builds
os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
os.environ.setdefault("OMP_NUM_THREADS", "1")

import numpy as np
import scipy.linalg  # noqa: F401  (loads SciPy's own OpenBLAS before the limit below)
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libttx.so")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def unpack_cores(g, shapes_key="core_shapes", data_key="core_data"):
    shapes = g[shapes_key]
    flat = g[data_key]
    out, off = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(flat[off:off + n].reshape(tuple(int(v) for v in s)))
        off += n
    return out


# The reference's AMEn trajectory depends on BLAS round-off (its own C1 run
# changes ranks with the OpenBLAS thread count), so the oracle runs with one
# BLAS thread, the setting of the reference's CI.
try:
    from threadpoolctl import threadpool_limits

    _BLAS_LIMIT = threadpool_limits(limits=1, user_api="blas")
except Exception:  # pragma: no cover
    _BLAS_LIMIT = None
