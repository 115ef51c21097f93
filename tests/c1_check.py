"""End-to-end checks of a C1 AMEn solution (2-D Poisson 256 x 256, 16 cores).

Shared by the per-kernel and engine tests.
"""

import numpy as np

from conftest import unpack_cores
from oracle import tt_oracle as O


def check_c1_solution(x_full, info, g):
    """C1 end to end against the reference's golden run and the exact
    discrete solution.

    * stop: converged, FP64 residual <= 1e-8 (no slack);
    * sweeps: the reference takes 6, and its sweep-5 residual (1.34e-8) is
      within 34% of the tolerance, while its per-sweep residuals move by up
      to 3x between equivalent runs (sweep 3: 9.6e-5 with LU, 3.2e-5 with
      Cholesky local solves; other BLAS thread counts: bond 33, or
      'stagnated').  The device's trajectory differs from the reference's by
      the frames of numerically rank-deficient enriched cores (every site
      decision on the reference's own states is identical:
      test_site_parity_gpu.py), so the sweep count is bounded to 6 +- 1 and
      the largest bond to 32 +- 1;
    * solution: within ||A x - b|| / lam_min of the exact discrete solution
      (DST closed form, tests/poisson_exact.py), and the reference's solution
      within its own bound, so the two differ by at most the sum.
    """
    import poisson_exact as PE
    assert str(info.stop_reason) == "converged"
    assert info.residuals[-1] <= 1e-8
    assert abs(info.num_sweeps_run - int(g["num_sweeps_run"])) <= 1
    b_full = np.ones(2**16)
    err, bound, nrm = PE.error_bound(x_full, b_full, 8)
    ref = O.Train(unpack_cores(g)).full()
    err_ref, bound_ref, _ = PE.error_bound(ref, b_full, 8)
    dist = float(np.linalg.norm(x_full.reshape(-1) - ref.reshape(-1)))
    print(f"C1: sweeps {info.num_sweeps_run} residual {info.residuals[-1]:.3e}; "
          f"|x - x_exact| {err:.3e} (bound {bound:.3e}); |x - x_ref|/|x_ref| {dist / nrm:.3e}",
          flush=True)
    assert err <= bound
    assert dist <= bound + bound_ref
    # The bound allows ~2e-8 relative; both solutions are in fact within
    # ~1e-11 of the exact one (their residuals are dominated by components
    # that barely move the solution), so the 1e-10 relative agreement the
    # drop-in asks for holds even though the trajectories differ (measured:
    # 5.4e-12 per-kernel, 3.6e-12 engine).
    assert dist / nrm <= 1e-10
    return dist / nrm
