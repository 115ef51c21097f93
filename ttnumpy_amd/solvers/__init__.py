"""Device-resident TT linear solvers (AMEn, ALS, MALS) with the reference vocabulary."""

from .common import (LocalSolver, SolverInfo, StopReason, TruncationMode,  # noqa: F401
                     relative_residual)
from .amen import amen  # noqa: F401
from .als import als  # noqa: F401
from .mals import mals  # noqa: F401
