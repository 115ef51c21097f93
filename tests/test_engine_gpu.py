"""The batched engine against the oracle, and the same request alone vs in a mix.

Every case prints its name before it runs.
"""

import numpy as np
import pytest

from conftest import load_golden, unpack_cores

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import tt_oracle as O  # noqa: E402
from ttnumpy_amd import problems as P  # noqa: E402


def _tt(cores):
    import ttnumpy_amd as tx
    return tx.TensorTrain([np.asarray(c) for c in cores])


def _poisson_request(n_bits, seed=0, x0_seed=0, **kw):
    import ttnumpy_amd as tx
    from ttnumpy_amd.engine import Request
    d = 2 * n_bits
    opts = dict(symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=seed)
    opts.update(kw)
    return Request(A=_tt(P.poisson_2d_cores(n_bits)), b=_tt(P.ones_cores([2] * d)),
                   x0=tx.rand([2] * d, ranks=2, seed=x0_seed), **opts)


def _rel_diff(x, ref: O.Train):
    xo = O.Train(x.to_numpy())
    return O.tt_norm(O.tt_add(xo, O.tt_scale(ref, -1.0))) / O.tt_norm(ref)


def test_engine_poisson16_matches_oracle():
    """16 x 16 grid (d = 8): direct path only; converged, same sweeps and bond
    profile as the oracle, solution within the residual-implied distance.

    The sweep-1 residual is only checked to a factor of 2: the enriched cores are
    nearly rank deficient (sigma_min / sigma_1 ~ 1e-7), so AMEn's first sweep
    follows the round-off of the factorizations.  The oracle itself moves from
    1.127e-3 to 1.105e-3 when LAPACK's QR is swapped for an equivalent
    Householder QR; on the device the same step gives 1.24e-3 with the LU local
    solve and 8.9e-4 with the Cholesky one (same matrices, different rounding).
    """
    from ttnumpy_amd.engine import amen_batched
    print("engine poisson 16x16", flush=True)
    req = _poisson_request(4)
    (out,) = amen_batched([req], rmax=32, poll=4)
    assert out.error is None, out.error
    A = O.Train(P.poisson_2d_cores(4))
    b = O.Train(P.ones_cores([2] * 8))
    x0 = O.Train(P.random_train_cores([2] * 8, 2, 0))
    xo, io = O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0)
    print(out.info, io, flush=True)
    assert str(out.info.stop_reason) == "converged"
    assert out.info.residuals[-1] <= 1e-8
    assert out.info.num_sweeps_run == io.num_sweeps_run
    assert out.info.max_bonds == io.max_bonds
    assert 0.5 * io.residuals[0] < out.info.residuals[0] < 2.0 * io.residuals[0]
    assert _rel_diff(out.result, xo) < 1e-7


def test_engine_equals_device_path_bitwise():
    """The engine re-hosts the per-kernel device path: identical residual histories and cores."""
    import ttnumpy_amd as tx
    from ttnumpy_amd.engine import amen_batched
    print("engine vs device path, poisson 64x64", flush=True)
    req = _poisson_request(6)
    (out,) = amen_batched([req], rmax=40, poll=8)
    x, info = tx.amen(req.A, req.b, x0=req.x0, symmetric=True, residual_tol=1e-8, max_sweeps=20,
                      seed=0, return_info=True)
    assert out.info.residuals == info.residuals
    assert out.result.bonds == x.bonds
    for a, b in zip(out.result.to_numpy(), x.to_numpy()):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-13 * max(1.0, np.abs(b).max()))


def test_engine_poisson64_minres_path():
    """64 x 64 grid with direct_solve_size=16: the MINRES path of the engine."""
    from ttnumpy_amd.engine import amen_batched
    print("engine poisson 64x64 minres", flush=True)
    g = load_golden("poisson64_amen_minres")
    (out,) = amen_batched([_poisson_request(6, direct_solve_size=16)], rmax=40, poll=8)
    assert out.error is None, out.error
    print(out.info, flush=True)
    assert str(out.info.stop_reason) == "converged"
    assert out.info.residuals[-1] <= 1e-8
    assert out.info.local_solves > 0
    assert _rel_diff(out.result, O.Train(unpack_cores(g))) < 1e-6


def test_engine_mixed_minres_path():
    """Mixed precision on the MINRES path (64 x 64 Poisson, direct_solve_size=16:
    every local solve iterative): the inner MINRES runs on the FP32 operator
    (FP32 MFMA contractions), the outer FP64 correction meets the reference's
    acceptance test ||A u - b|| <= local_rtol ||b||, so the request converges
    to the FP64 tolerance (FP64-verified) and lands on the FP64 request's
    solution within what the two residuals allow."""
    import poisson_exact as PE
    from ttnumpy_amd.engine import Engine
    from ttnumpy_amd.solvers.common import relative_residual
    print("engine mixed minres path", flush=True)
    eng = Engine(2, rmax=40)
    r64 = _poisson_request(6, direct_solve_size=16)
    r32 = _poisson_request(6, direct_solve_size=16, precision="mixed")
    eng.run([r64, r32], poll=8)
    assert r64.error is None and r32.error is None, (r64.error, r32.error)
    print(r64.info, r32.info, eng.mixed_solves, eng.mixed_refines, flush=True)
    assert eng.mixed_solves >= r32.info.local_solves > 0
    assert r32.info.local_inexact == 0
    assert str(r32.info.stop_reason) == "converged"
    res = float(relative_residual(r32.A, r32.b, r32.result))
    assert res <= 1e-8
    d = _rel_diff(r32.result, O.Train(r64.result.to_numpy()))
    # both within ||A x - b|| / lam_min of the exact solution; relative to
    # ||x||: cond(A) (r_32 + r_64)
    bound = PE.kappa(6) * (res + r64.info.residuals[-1])
    print("relative difference mixed vs fp64", d, "bound", bound, "fp64-verified residual", res,
          flush=True)
    assert d <= bound


def test_engine_bf16_minres_c5_shape():
    """BF16 operands on the MINRES path (precision="bf16"): the C5 system at
    d = 6 (A = I + sum_i L_i, kappa(A) < 30, mode 8), direct_solve_size 64 so
    that most local solves are iterative.  The inner MINRES applies the local
    operator with its three products on BF16 operands and FP32 accumulation
    (v_mfma_f32_16x16x32_bf16), the outer FP64 correction meets the reference's
    acceptance test, so the request ends like the FP64 one (residual <= 3e-8,
    FP64-verified) and its solution is within what the two residuals allow of
    the FP64 request's: kappa(A) (r_bf16 + r_64), kappa(A) < 30."""
    from ttnumpy_amd.engine import Engine, Request
    from ttnumpy_amd.solvers.common import relative_residual
    print("engine bf16 minres C5 shape d=6", flush=True)
    A, b = _tt(P.shifted_laplacian_sum_cores(6, 8, 1.0 / 6)), _tt(P.random_train_cores([8] * 6, 4, 3))
    kw = dict(max_rank=64, residual_tol=1e-8, max_sweeps=20, symmetric=True, seed=0,
              direct_solve_size=64)
    reqs = {p: Request(A=A, b=b, precision=p, **kw) for p in ("fp64", "mixed", "bf16")}
    out = {}
    for p, r in reqs.items():  # one engine each: the counters are per engine
        eng = Engine(1, rmax=64)
        eng.run([r], poll=8)
        assert r.error is None, (p, r.error)
        out[p] = (r, eng.minres_iters, eng.mixed_solves, eng.mixed_refines)
        print(p, r.info, "minres iterations", eng.minres_iters, "mixed solves", eng.mixed_solves,
              "outer passes", eng.mixed_refines, flush=True)
    r64 = out["fp64"][0]
    for p in ("mixed", "bf16"):
        r = out[p][0]
        assert out[p][2] > 0  # mixed solves ran
        assert str(r.info.stop_reason) in ("converged", "stagnated")
        res = float(relative_residual(A, b, r.result))
        assert res <= 3e-8, (p, res)
        d = _rel_diff(r.result, O.Train(r64.result.to_numpy()))
        print(p, "relative difference vs fp64", d, "fp64-verified residual", res, flush=True)
        assert d <= 30.0 * (res + r64.info.residuals[-1]), (p, d)


def test_engine_dense_fallback_matches_oracle():
    """allow_dense_fallback (common.py:793-809): when the four MINRES attempts
    miss local_rtol (here 1e-15, below what MINRES reaches), the engine solves
    the explicit local matrix instead, as the reference does; the solve
    converges, counts its fallbacks, and lands on the oracle's solution
    (which takes the same fallbacks) within the residual bound."""
    import poisson_exact as PE
    from ttnumpy_amd.engine import Engine
    print("engine dense fallback", flush=True)
    kw = dict(direct_solve_size=16, local_rtol=1e-15, allow_dense_fallback=True)
    req = _poisson_request(6, **kw)
    Engine(1, rmax=40).run([req], poll=8)
    assert req.error is None, req.error
    A = O.Train(P.poisson_2d_cores(6))
    b = O.Train(P.ones_cores([2] * 12))
    x0 = O.Train(P.random_train_cores([2] * 12, 2, 0))
    xo, io = O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0, **kw)
    print(req.info, io, flush=True)
    assert req.info.local_fallbacks > 0 and io.local_fallbacks > 0
    assert req.info.local_inexact == 0 and io.local_inexact == 0
    assert str(req.info.stop_reason) == str(io.stop_reason) == "converged"
    assert req.info.residuals[-1] <= 1e-8
    d = _rel_diff(req.result, xo)
    bound = PE.kappa(6) * (req.info.residuals[-1] + io.residuals[-1])
    print("relative difference vs oracle", d, "bound", bound, flush=True)
    assert d <= bound


def test_engine_gmres_nonsymmetric_matches_device_path_and_oracle():
    """symmetric=False with iterative local solves (convection-diffusion, the
    operator of test_api_gpu's GMRES test): the engine runs SciPy's restarted
    GMRES (gmres_blk.h, the routines of the per-kernel path's GMRES kernels)
    with the reference's four attempts (common.py:776-809).  It converges as
    the per-kernel path and the oracle (scipy's gmres) do, and all three land
    within the residual-implied distance of the dense solution."""
    import ttnumpy_amd as tx
    from test_api_gpu import _convection_diffusion_cores
    from ttnumpy_amd.engine import Engine, Request
    print("engine gmres nonsymmetric", flush=True)
    L = 6
    A_c = _convection_diffusion_cores(L, 20.0)
    b_c = P.ones_cores([2] * L)
    kw = dict(residual_tol=1e-8, max_sweeps=10, symmetric=False, direct_solve_size=4, max_rank=16,
              seed=0)
    x0 = tx.rand([2] * L, ranks=2, seed=0)
    req = Request(A=_tt(A_c), b=_tt(b_c), x0=x0, **kw)
    eng = Engine(1, rmax=32)
    eng.run([req], poll=8)
    assert req.error is None, req.error
    x, info = tx.amen(_tt(A_c), _tt(b_c), x0=x0, return_info=True, **kw)
    xo, io = O.amen(O.Train(A_c), O.Train(b_c), O.Train(x0.to_numpy()), **kw)
    print(req.info, info, io, eng.minres_iters, flush=True)
    Ad = O.Train(A_c).dense_matrix()
    ref = np.linalg.solve(Ad, np.ones(2**L))
    cond = np.linalg.cond(Ad)
    assert req.info.local_solves > 0 and eng.minres_iters > 0
    for stop, res in ((req.info.stop_reason, req.info.residuals), (info.stop_reason, info.residuals),
                      (io.stop_reason, io.residuals)):
        assert str(stop) in ("converged", "stagnated")
        assert res[-1] <= 3e-8
    assert abs(req.info.num_sweeps_run - io.num_sweeps_run) <= 2
    assert abs(req.info.num_sweeps_run - info.num_sweeps_run) <= 2
    xe = req.result.get_full_tensor().reshape(-1)
    assert np.linalg.norm(xe - ref) <= 3e-8 * cond * np.linalg.norm(ref)


def test_engine_c1_poisson256():
    """C1 on the engine: the checks of c1_check.check_c1_solution
    (converged at 1e-8, 6 +- 1 sweeps, the exact discrete solution within the
    residual bound)."""
    from c1_check import check_c1_solution
    from ttnumpy_amd.engine import amen_batched
    print("engine C1 poisson 256x256", flush=True)
    g = load_golden("c1_poisson256_amen")
    (out,) = amen_batched([_poisson_request(8)], rmax=48, poll=16)
    assert out.error is None, out.error
    print(out.info, out.result.bonds, flush=True)
    check_c1_solution(out.result.get_full_tensor(), out.info, g)
    assert abs(max(out.result.bonds) - max(g["bonds"])) <= 1


def test_engine_rank_capacity_grows():
    """A bond that outgrows the slot's rank capacity re-plans the slot with
    twice the capacity (the reference bounds ranks only by max_rank,
    amen.py:441-451) and serves the request again from its x0; the result is
    the one a slot of that capacity gives from the start, and the oracle's."""
    from ttnumpy_amd.engine import Engine
    print("engine rank capacity growth", flush=True)
    g = load_golden("poisson64_amen")
    small = Engine(2, rmax=8)
    r1 = _poisson_request(6)                 # max bond 24
    r2 = _poisson_request(5, seed=3, max_rank=10)
    small.run([r1, r2], poll=8)
    print(small.regrown, r1.rank_capacity, r2.rank_capacity, r1.info, flush=True)
    assert r1.error is None and r2.error is None, (r1.error, r2.error)
    assert small.regrown >= 3 and r1.rank_capacity == 32 and r2.rank_capacity == 10
    assert max(r2.result.bonds) <= 10
    ref = _poisson_request(6)
    Engine(1, rmax=32).run([ref], poll=8)
    assert r1.info.residuals == ref.info.residuals and r1.result.bonds == ref.result.bonds
    for a, b in zip(r1.result.to_numpy(), ref.result.to_numpy()):
        np.testing.assert_array_equal(a, b)
    assert str(r1.info.stop_reason) == "converged"
    assert r1.info.num_sweeps_run == int(g["num_sweeps_run"])
    assert _rel_diff(r1.result, O.Train(unpack_cores(g))) < 1e-7


def test_engine_alone_equals_in_mix():
    """A request's result is bit-identical alone and batched with different requests."""
    from ttnumpy_amd.engine import Engine
    print("engine alone vs mix", flush=True)
    alone = Engine(1, rmax=40).run([_poisson_request(5, seed=1, x0_seed=1)], poll=4)[0]
    mix_reqs = [_poisson_request(4, seed=2), _poisson_request(5, seed=1, x0_seed=1),
                _poisson_request(6, seed=3, x0_seed=5), _poisson_request(5, seed=7, x0_seed=2)]
    mix = Engine(3, rmax=40).run(mix_reqs, poll=4)
    same = [r for r in mix if r is mix_reqs[1]][0]
    assert alone.error is None and same.error is None
    assert alone.info.residuals == same.info.residuals
    assert alone.result.bonds == same.result.bonds
    for a, b in zip(alone.result.to_numpy(), same.result.to_numpy()):
        np.testing.assert_array_equal(a, b)
    for r in mix:
        assert r.error is None and r.info.residuals[-1] <= 1e-8


def test_serving_diffusion_alone_equals_in_stream():
    """Parametric diffusion requests of different grid sizes served as a timed
    stream (one engine per level on its own stream, continuous admission,
    MINRES slices over ticks, paced polls of the fused stages) give the same
    bits as the same request served alone on the seven tick kernels."""
    from ttnumpy_amd.engine import Engine
    from ttnumpy_amd.serving import ServingConfig, ServingItem, build_requests, serve
    print("serving diffusion alone vs stream", flush=True)
    cfg = ServingConfig(n_requests=6, levels=(4, 5, 6), rate=200.0, nslots=3, rmax=40, poll=4,
                        seed=11, per_level=True)
    rep, done = serve(cfg)
    print(rep, flush=True)
    assert rep.errors == 0 and rep.solves == 6
    target = done[0]
    it = target.tag
    (alone,) = build_requests(cfg, [ServingItem(it.index, it.level, it.theta, 0.0)])
    Engine(1, rmax=40).run([alone], poll=4)
    assert alone.error is None
    assert alone.info.residuals == target.info.residuals
    assert alone.result.bonds == target.result.bonds
    for a, b in zip(alone.result.to_numpy(), target.result.to_numpy()):
        np.testing.assert_array_equal(a, b)


def test_serving_c4_stream_matches_oracle():
    """C4's parametric diffusion stream (``examples/parametric_diffusion.py:35-58``:
    k = 1 + theta sin(pi x) sin(pi y), b = ones), here at L = 6..8, served with
    continuous admission and paced polls by one engine for all levels, as the
    serving loop does by default.  Each result is
    compared with the oracle's AMEn on the same operator, right-hand side,
    initial guess and seed, and with the exact discrete solution (sparse
    direct solve): the same stop reason, sweeps within one, the error within
    ||A x - b|| / lam_min for both, and the two solutions within 1e-9 relative
    of each other."""
    import poisson_exact as PE
    from ttnumpy_amd.serving import ServingConfig, serve
    print("serving C4 stream vs oracle", flush=True)
    cfg = ServingConfig(n_requests=9, levels=(6, 7, 8), rate=100.0, nslots=3, rmax=64, poll=8,
                        seed=5)
    rep, done = serve(cfg)
    print(rep, flush=True)
    assert rep.errors == 0 and rep.solves == 9
    worst = 0.0
    for r in sorted(done, key=lambda r: r.tag.index):
        it = r.tag
        L = it.level
        A = O.Train(P.diffusion_2d_cores(L, it.theta))
        b = O.Train(P.ones_cores([2] * (2 * L)))
        x0 = O.Train(P.random_train_cores([2] * (2 * L), 2, it.index))
        xo, io = O.amen(A, b, x0, symmetric=True, residual_tol=cfg.residual_tol,
                        max_sweeps=cfg.max_sweeps, seed=it.index)
        xg = r.result.get_full_tensor().reshape(-1)
        err, bound, nrm = PE.diffusion_error_bound(xg, L, it.theta)
        err_o, bound_o, _ = PE.diffusion_error_bound(xo.full(), L, it.theta)
        dist = float(np.linalg.norm(xg - xo.full().reshape(-1))) / nrm
        worst = max(worst, dist)
        print(f"  request {it.index} L={L} theta={it.theta:.3f}: gpu {r.info.stop_reason} "
              f"{r.info.num_sweeps_run} sweeps res {r.info.residuals[-1]:.2e} bond "
              f"{max(r.result.bonds)}; oracle {io.stop_reason} {io.num_sweeps_run} sweeps res "
              f"{io.residuals[-1]:.2e} bond {max(xo.bonds)}; |x-x_exact| {err:.2e} (bound "
              f"{bound:.2e}); |x_gpu-x_oracle|/|x| {dist:.2e}", flush=True)
        assert str(r.info.stop_reason) == str(io.stop_reason)
        if str(io.stop_reason) == "converged":
            assert r.info.residuals[-1] <= cfg.residual_tol
        assert abs(r.info.num_sweeps_run - io.num_sweeps_run) <= 1
        assert err <= bound and err_o <= bound_o
        assert dist * nrm <= bound + bound_o
    print(f"worst relative distance gpu vs oracle {worst:.2e}", flush=True)
    assert worst <= 1e-9


def test_engine_mixed_precision_c1():
    """Mixed precision (FP32 Cholesky + FP64 refinement for direct solves of
    size >= 256): converges to the FP64 tolerance, FP64-verified, and lands on
    the FP64 engine's solution within the residual-implied distance."""
    from ttnumpy_amd.engine import Engine
    from ttnumpy_amd.solvers.common import relative_residual
    print("engine C1 mixed vs fp64", flush=True)
    eng = Engine(2, rmax=48)
    r64 = _poisson_request(8)
    r32 = _poisson_request(8, precision="mixed")
    out = eng.run([r64, r32], poll=16)
    assert all(r.error is None for r in out)
    print(r64.info, r32.info, eng.mixed_solves, eng.mixed_refines, eng.mixed_fallbacks, flush=True)
    assert eng.mixed_solves > 0 and eng.mixed_fallbacks == 0
    assert str(r32.info.stop_reason) == "converged"
    res = float(relative_residual(r32.A, r32.b, r32.result))
    assert res <= 1e-8
    d = _rel_diff(r32.result, O.Train(r64.result.to_numpy()))
    print("relative difference mixed vs fp64", d, "fp64-verified residual", res, flush=True)
    # measured 6.5e-12: both land on the exact discrete solution to ~1e-11
    # (the residual bound ||A x - b|| / lam_min allows ~1e-8 relative)
    assert d < 1e-10


def test_engine_graph_replays_equal_eager_ticks():
    """Polls replayed from one recorded HIP graph give bit-identical results
    to eager tick launches: a mix of requests admitted while others run (the
    slot table changes between replays, with host work between them), served
    twice on the same graph engine."""
    from ttnumpy_amd.engine import Engine
    print("engine graph replays vs eager", flush=True)

    def mix():
        return [_poisson_request(4, seed=2), _poisson_request(5, seed=1, x0_seed=1),
                _poisson_request(6, seed=3, x0_seed=5), _poisson_request(5, seed=7, x0_seed=2)]

    eager = Engine(3, rmax=40, graphs=False).run(mix(), poll=4)
    geng = Engine(3, rmax=40, graphs=True)
    for rep in range(2):
        graph = geng.run(mix(), poll=4)
        assert len(geng._graphs) == 1
        for a, b in zip(sorted(eager, key=lambda r: (r.seed, len(r.A))),
                        sorted(graph, key=lambda r: (r.seed, len(r.A)))):
            assert a.error is None and b.error is None
            assert a.info.residuals == b.info.residuals
            assert a.result.bonds == b.result.bonds
            for ca, cb in zip(a.result.to_numpy(), b.result.to_numpy()):
                np.testing.assert_array_equal(ca, cb)
    geng.close_graphs()


def test_engine_idle_stage_skipping_and_fused_bitwise_equal():
    """Ticks that leave out the FP32 and LU kernels while no resident slot
    needs them, ticks as one fused kernel (each slot runs the seven stages
    without waiting for the others), and paced polls (each slot runs ticks at
    its own pace for a wall-clock budget) give the same bits as ticks that
    launch all seven kernels.  Three slots serve a queue whose mix changes as
    requests retire: symmetric FP64 requests, a negated Poisson operator
    (negative definite: every Cholesky fails and the solve turns to LU
    between polls, then waits for the next poll's LU kernel), a
    mixed-precision request and a nonsymmetric GMRES request."""
    import ttnumpy_amd as tx
    from test_api_gpu import _convection_diffusion_cores
    from ttnumpy_amd.engine import FUSED, STAGES_ALL, STAGE_CHOL32, STAGE_LU, Engine, Request
    print("engine idle stage skipping", flush=True)

    def queue():
        neg_a, neg_b = P.poisson_2d_cores(4), P.ones_cores([2] * 8)
        neg_a[0], neg_b[0] = -neg_a[0], -neg_b[0]
        cd = _convection_diffusion_cores(6, 20.0)
        return [_poisson_request(4, seed=2),
                Request(A=_tt(neg_a), b=_tt(neg_b), x0=tx.rand([2] * 8, ranks=2, seed=3),
                        symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=3),
                _poisson_request(5, seed=1, x0_seed=1),
                _poisson_request(5, seed=4, x0_seed=4, precision="mixed"),
                Request(A=_tt(cd), b=_tt(P.ones_cores([2] * 6)), x0=tx.rand([2] * 6, ranks=2, seed=0),
                        symmetric=False, residual_tol=1e-8, max_sweeps=10, direct_solve_size=4,
                        max_rank=16, seed=0),
                _poisson_request(4, seed=5, x0_seed=6)]

    out = {}
    for mode in ("all", "skip", "fused", "paced"):
        eng = Engine(3, rmax=40, fused=mode in ("fused", "paced"))
        eng.skip_idle_stages = mode == "skip"
        eng.pace_us = 3000.0 if mode == "paced" else 0.0
        done = sorted(eng.run(queue(), poll=4), key=lambda r: r.seed)
        out[mode] = (done, {k[1] for k in eng._graphs}, eng.chol_fallbacks, eng.mixed_solves,
                     eng.paced_polls)
        eng.close_graphs()
    full, full_masks, full_lu, full_mixed, _ = out["all"]
    print({m: (sorted(o[1]), o[2], o[3], o[4]) for m, o in out.items()}, flush=True)
    assert full_masks == {STAGES_ALL}
    # the paced engine serves polls with one slot left (or slots in step) by
    # lockstep ticks of the seven kernels, the others by paced launches
    assert out["fused"][1] == {FUSED} and FUSED not in out["paced"][1]
    assert out["paced"][4] > 0 and out["fused"][4] == 0
    skip_masks = out["skip"][1]
    assert STAGES_ALL & ~(STAGE_CHOL32 | STAGE_LU) in skip_masks and len(skip_masks) > 1
    assert full_lu > 0
    for mode in ("skip", "fused", "paced"):
        other, _, lu, mixed, _ = out[mode]
        assert lu == full_lu and mixed == full_mixed
        for a, b in zip(full, other):
            assert a.seed == b.seed and a.error is None and b.error is None, (mode, a.error, b.error)
            assert a.info.residuals == b.info.residuals
            assert a.info.num_sweeps_run == b.info.num_sweeps_run
            assert a.result.bonds == b.result.bonds
            for ca, cb in zip(a.result.to_numpy(), b.result.to_numpy()):
                np.testing.assert_array_equal(ca, cb)
    assert all(str(r.info.stop_reason) in ("converged", "stagnated") for r in full)


def test_engine_paced_aligned_polls_use_tick_kernels():
    """A paced engine serves polls whose running slots all hold the same
    request at the same step with lockstep ticks of the seven kernels (tick
    graphs get recorded, no paced launch), and polls of a mixed queue with
    paced launches; either way the results are the bits of a lockstep
    engine."""
    import dataclasses
    from ttnumpy_amd.engine import STAGES_ALL, STAGE_CHOL32, STAGE_LU, Engine
    print("engine paced aligned polls", flush=True)
    one = _poisson_request(5, seed=1, x0_seed=1)
    same = [dataclasses.replace(one) for _ in range(3)]  # the same trains: one descriptor
    ref = Engine(3, rmax=40).run([_poisson_request(5, seed=1, x0_seed=1)], poll=4)[0]
    eng = Engine(3, rmax=40, fused=True)
    eng.pace_us = 2000.0
    eng.run(same, poll=4)
    assert {k[1] for k in eng._graphs} == {STAGES_ALL & ~(STAGE_CHOL32 | STAGE_LU)}
    assert eng.paced_polls == 0
    mix = [_poisson_request(4, seed=2), _poisson_request(5, seed=1, x0_seed=1),
           _poisson_request(6, seed=3, x0_seed=5)]
    eng2 = Engine(3, rmax=40, fused=True)
    eng2.pace_us = 2000.0
    eng2.run(mix, poll=4)
    assert eng2.paced_polls > 0
    for r in same + [mix[1]]:
        assert r.error is None
        assert r.info.residuals == ref.info.residuals
        assert r.result.bonds == ref.result.bonds
        for a, b in zip(r.result.to_numpy(), ref.result.to_numpy()):
            np.testing.assert_array_equal(a, b)
    eng.close_graphs()


def test_engine_tile_variants_bitwise_equal_and_tuner():
    """The Cholesky tile variants (128- and 64-row tiles below the diagonal
    blocks) sum every element over the same k in the same order, so a mix with direct
    and MINRES local solves gives identical residual histories and cores under
    either; the engine tuner then picks one of them from measured kernel time."""
    from ttnumpy_amd.autotune import ENGINE_TILE_VARIANTS, tune_engine
    from ttnumpy_amd.engine import Engine
    print("engine tile variants, poisson 32x32 and 64x64", flush=True)

    def mix():
        return [_poisson_request(6, direct_solve_size=600), _poisson_request(5, seed=1, x0_seed=1)]

    runs = []  # local_solves counts the iterative ones: the 64 x 64 case has both kinds
    for t in sorted(ENGINE_TILE_VARIANTS):
        eng = Engine(2, rmax=40)
        eng.tiles = t
        runs.append(sorted(eng.run(mix(), poll=8), key=lambda r: len(r.A)))
    assert runs[0][-1].info.local_solves > 0
    for a, b in zip(*runs):
        assert a.error is None and b.error is None
        assert a.info.residuals == b.info.residuals
        assert a.result.bonds == b.result.bonds
        for ca, cb in zip(a.result.to_numpy(), b.result.to_numpy()):
            np.testing.assert_array_equal(ca, cb)
    eng = Engine(2, rmax=40)
    rep = tune_engine(eng, mix(), poll=8)
    print(rep, flush=True)
    assert eng.tiles == rep["tiles"] and rep["tiles"] in ENGINE_TILE_VARIANTS
    assert all(rep["kernel_ms"][t]["total"] > 0 for t in ENGINE_TILE_VARIANTS)
