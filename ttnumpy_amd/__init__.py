"""MI355X-native tensor-train solver hot path with ttnumpy's Python API.

``amen``, ``als``, ``mals``, ``TensorTrain.truncate``, ``SolverInfo``/``StopReason``
mirror ttnumpy 0.3.0; all arithmetic runs in the HIP library ``libttx.so``
(built by :mod:`ttnumpy_amd.build`).  There is no CPU fallback.
"""

__version__ = "0.1.0"

from .tensor_train import (DimensionMismatch, TensorTrain, WrongTTFormat,  # noqa: F401
                           WrongTTType, eye, ones, rand, zeros)
from .solvers import (LocalSolver, SolverInfo, StopReason, TruncationMode, als,  # noqa: F401
                      amen, mals, relative_residual)
