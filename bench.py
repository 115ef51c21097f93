"""Benchmark: AMEn solves/s and latency on the 2D Poisson 256x256 QTT system, one MI355X.

Workload (C1): A = -Laplace on a 256 x 256 grid (Dirichlet, h = 1/257) as the
Kronecker-sum QTT operator of 16 cores with 2x2 modes and TTM bonds <= 4,
b = all-ones rank-1 train, x0 = rand([2]*16, ranks=2, seed=0);
amen(symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0) in FP64.

One step = one batch of ``--batch`` (default 512: two request slots per CU)
such solves served concurrently by the device-resident engine (one slot per
request, continuous batching); inputs
are resident in HBM before the clock starts.  ``value`` is solves per second
over the timed steps; p50/p99 are per-request latencies (admission to
result).  Prints one JSON line.  ``--gpus N`` with N > 1 exits: the benchmark
runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_BITS = 8  # 256 x 256 grid
RMAX = 48   # slot rank capacity (the C1 solution reaches bond 33)
PEAK_FP64_TFLOPS = 78.6
BATCH_PROFILED = 512  # batch of the committed counter summary
KERNELS = ("ttx::eng_env_kernel", "ttx::eng_solve_kernel", "ttx::eng_chol32_kernel",
           "ttx::eng_chol_kernel", "ttx::eng_lu_kernel", "ttx::eng_post_kernel",
           "ttx::eng_check_kernel")


def build_inputs():
    import ttnumpy_amd as tx
    from ttnumpy_amd import problems as P
    d = 2 * N_BITS
    A = tx.TensorTrain(P.poisson_2d_cores(N_BITS))
    b = tx.TensorTrain(P.ones_cores([2] * d))
    x0 = tx.rand([2] * d, ranks=2, seed=0)
    return A, b, x0


def requests(A, b, x0, n, precision="fp64"):
    from ttnumpy_amd.engine import Request
    return [Request(A=A, b=b, x0=x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0,
                    precision=precision) for _ in range(n)]


def timed_batches(eng, A, b, x0, args, precision):
    """``args.warmup`` untimed then ``args.steps`` timed batches; (solves/s, latencies, done, s)."""
    import torch
    for _ in range(args.warmup):
        eng.run(requests(A, b, x0, args.batch, precision), poll=args.poll)
    torch.cuda.synchronize()
    lat, done = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = eng.run(requests(A, b, x0, args.batch, precision), poll=args.poll)
        done += out
        lat += [r.finished - r.admitted for r in out]
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    ok = [r for r in done if r.error is None]
    return len(ok) / total, lat, done, total


def measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed FETCH_SIZE/WRITE_SIZE summary.

    The counters come from separate ``rocprofv3 --pmc`` runs of this benchmark
    (a counter run cannot time); the summary names the commit it measured.
    """
    short = kernel.split("::")[-1]
    path = os.path.join(ROOT, "profiles", f"{short}_c1_batch{BATCH_PROFILED}_traffic.json")
    if not os.path.exists(path):
        return None, None
    t = json.load(open(path))
    return t["bytes_per_launch"], {"file": os.path.relpath(path, ROOT), "commit": t.get("commit"),
                                   "read_bytes": t["read_bytes"], "write_bytes": t["write_bytes"]}


def verify_result(A, b, req):
    """FP64 check of one returned solution: ||A x - b|| / ||b|| in TT arithmetic on
    the GPU (the device relative_residual) and, on the CPU, by the oracle."""
    from ttnumpy_amd.solvers.common import relative_residual
    from oracle import tt_oracle as O
    from ttnumpy_amd import problems as P
    gpu = float(relative_residual(A, b, req.result))
    d = 2 * N_BITS
    cpu = O.relative_residual(O.Train(P.poisson_2d_cores(N_BITS)), O.Train(P.ones_cores([2] * d)),
                              O.Train(req.result.to_numpy()))
    return {"relative_residual_gpu_fp64": gpu, "relative_residual_cpu_oracle": float(cpu),
            "engine_reported": req.info.residuals[-1]}


def c2_rounding(with_cpu: bool):
    """C2: (x1 + x2).truncate(error=1e-10, max_rank=96) for two d=40, mode-4,
    rank-64 normal trains; FP64, mixed (FP32 absorb contractions) and
    randomized (method="randomized", seed 0) on the GPU, the oracle on one CPU
    thread."""
    import torch
    import ttnumpy_amd as tx
    from ttnumpy_amd import problems as P
    cores1 = P.random_train_cores([4] * 40, 64, 1, normal=True)
    cores2 = P.random_train_cores([4] * 40, 64, 2, normal=True)
    s = tx.TensorTrain(cores1).add(tx.TensorTrain(cores2))
    out = {}
    spectra = {}
    from ttnumpy_amd import device as D
    for prec in ("fp64", "mixed"):
        s.truncate(error=1e-10, max_rank=96, precision=prec)  # warm
        torch.cuda.synchronize()
        sp = []
        f0 = D.flops()
        t0 = time.perf_counter()
        o = s.truncate(error=1e-10, max_rank=96, precision=prec, spectra=sp)
        torch.cuda.synchronize()
        out[prec] = {"ms": 1e3 * (time.perf_counter() - t0), "bonds_max": max(o.bonds),
                     "gflop": (D.flops() - f0) / 1e9}
        spectra[prec] = ([x.cpu().numpy() for x in sp], o.bonds)
    b64 = spectra["fp64"][1]
    worst = 0.0
    for i, (a, b) in enumerate(zip(spectra["mixed"][0], spectra["fp64"][0])):
        kept = b64[len(b64) - 2 - i]
        worst = max(worst, float(np.max(np.abs(a[:kept] - b[:kept]) / b[:kept])))
    gf = out["fp64"]["gflop"]
    res = {"gpu_fp64_ms": out["fp64"]["ms"], "gpu_mixed_ms": out["mixed"]["ms"],
           "bonds_identical": spectra["mixed"][1] == b64, "max_bond": max(b64),
           "mixed_max_rel_singular_value_diff": worst, "gflop": gf,
           "gflop_per_s": gf / (1e-3 * out["fp64"]["ms"]),
           "frac_fp64_peak": gf / (1e-3 * out["fp64"]["ms"]) / (1e3 * PEAK_FP64_TFLOPS)}
    # randomized rounding (sketch bond 96 + 10 < the train's 128): time and its
    # distance to the direct rounding
    s.truncate(error=1e-10, max_rank=96, method="randomized", seed=0)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rnd = s.truncate(error=1e-10, max_rank=96, method="randomized", seed=0)
    torch.cuda.synchronize()
    res["gpu_randomized_ms"] = 1e3 * (time.perf_counter() - t0)
    direct = s.truncate(error=1e-10, max_rank=96)
    res["randomized_rel_diff_vs_direct"] = float(
        (rnd - direct).tt_norm() / direct.tt_norm())
    if with_cpu:
        from threadpoolctl import threadpool_limits
        from oracle import tt_oracle as O
        so = O.tt_add(O.Train(cores1), O.Train(cores2))
        with threadpool_limits(1):
            t0 = time.perf_counter()
            O.truncate(so, error=1e-10, max_rank=96)
            res["cpu_oracle_1thread_ms"] = 1e3 * (time.perf_counter() - t0)
            t0 = time.perf_counter()
            O.truncate(so, error=1e-10, max_rank=96, method="randomized", seed=0)
            res["cpu_oracle_randomized_1thread_ms"] = 1e3 * (time.perf_counter() - t0)
    return res


def c5_high_dim(with_cpu: bool):
    """C5: AMEn on A = I + sum_i L_i (d=64, mode 8, rank-2 TTM, L = tridiag(-1,2,-1)),
    b = rand([8]*64, ranks=16, seed=3), x0 = ones with b's bonds, max_rank=128,
    residual_tol=1e-8, max_sweeps=30, symmetric, seed 0; the per-kernel device
    path (MINRES local solves, multi-workgroup MFMA contractions), then the same
    solve on the batched engine.  Also the per-kernel run with the stagnation stop
    off (30 sweeps), which drives the ranks up."""
    import torch
    import ttnumpy_amd as tx
    from ttnumpy_amd import device as D
    from ttnumpy_amd import problems as P
    A_c = P.shifted_laplacian_sum_cores(64, 8, 1.0 / 64)
    b_c = P.random_train_cores([8] * 64, 16, 3)
    A, b = tx.TensorTrain(A_c), tx.TensorTrain(b_c)
    res = {}
    # one untimed run first (kernel loading and allocator growth for the new
    # shapes), recorded by the GEMM-tiling autotuner, which then sets the tiling
    from ttnumpy_amd import autotune
    res["gemm_tiling"] = autotune.tune(lambda: tx.amen(A, b, max_rank=128, residual_tol=1e-8,
                                                       max_sweeps=30, symmetric=True, seed=0))
    for name, stag in (("auto", "auto"), ("no_stagnation_stop", None)):
        torch.cuda.synchronize()
        f0 = D.flops()
        t0 = time.perf_counter()
        x, info = tx.amen(A, b, max_rank=128, residual_tol=1e-8, max_sweeps=30, symmetric=True,
                          seed=0, return_info=True, stagnation_factor=stag)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gf = (D.flops() - f0) / 1e9
        res[name] = {"gpu_s": wall, "stop_reason": str(info.stop_reason),
                     "sweeps": info.num_sweeps_run, "max_bond": max(x.bonds),
                     "residual_fp64": info.residuals[-1], "local_solves": info.local_solves,
                     "gflop": gf, "gflop_per_s": gf / wall,
                     "frac_fp64_peak": gf / wall / (1e3 * PEAK_FP64_TFLOPS)}
    # the same solve on the batched engine (one slot, rank capacity 64): every
    # site step of the sweep stays on the device between the host's polls
    from ttnumpy_amd.engine import Engine, Request
    eng = Engine(1, rmax=64)
    for rep in range(2):          # the first run includes the slot set-up
        req = Request(A=A, b=b, max_rank=128, residual_tol=1e-8, max_sweeps=30, symmetric=True,
                      seed=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = eng.run([req])[0]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    if r.error is None:
        res["engine"] = {"gpu_s": wall, "stop_reason": str(r.info.stop_reason),
                         "sweeps": r.info.num_sweeps_run, "max_bond": max(r.result.bonds),
                         "residual_fp64": r.info.residuals[-1], "local_solves": r.info.local_solves,
                         "rmax": 64}
    else:
        res["engine"] = {"error": r.error}
    # the mixed form on the same engine: MINRES local solves on the FP32 operator
    # (FP32 MFMA contractions) with FP64 defect correction, FP32 Cholesky with FP64
    # refinement for the direct ones; residual check, QR/SVD and rank rule in FP64.
    # "bf16": the same with the MINRES operator's products on BF16 operands (FP32
    # accumulation, v_mfma_f32_16x16x32_bf16)
    from ttnumpy_amd.solvers.common import relative_residual
    for prec in ("mixed", "bf16"):
        engm = Engine(1, rmax=64)
        for rep in range(2):
            reqm = Request(A=A, b=b, max_rank=128, residual_tol=1e-8, max_sweeps=30,
                           symmetric=True, seed=0, precision=prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rm = engm.run([reqm])[0]
            torch.cuda.synchronize()
            wallm = time.perf_counter() - t0
        if rm.error is None and r.error is None:
            from oracle import tt_oracle as O
            x64, xm = O.Train(r.result.to_numpy()), O.Train(rm.result.to_numpy())
            res[f"engine_{prec}"] = {
                "gpu_s": wallm, "ratio_vs_fp64": wall / wallm,
                "stop_reason": str(rm.info.stop_reason),
                "sweeps": rm.info.num_sweeps_run, "max_bond": max(rm.result.bonds),
                "residual_fp64_verified": float(relative_residual(A, b, rm.result)),
                "residual_fp64_verified_fp64_path": float(relative_residual(A, b, r.result)),
                "relative_difference_vs_fp64_solution": float(
                    O.tt_norm(O.tt_add(xm, O.tt_scale(x64, -1.0))) / O.tt_norm(x64)),
                "local_solves": rm.info.local_solves, "local_inexact": rm.info.local_inexact,
                "minres_iters": engm.minres_iters // 2,
                "outer_corrections": engm.mixed_refines // 2,
                "minres_iters_fp64": eng.minres_iters // 2}
        else:
            res[f"engine_{prec}"] = {"error": rm.error or r.error}
    if with_cpu:
        from threadpoolctl import threadpool_limits
        from oracle import tt_oracle as O
        with threadpool_limits(1):
            t0 = time.perf_counter()
            xo, io = O.amen(O.Train(A_c), O.Train(b_c), max_rank=128, residual_tol=1e-8,
                            max_sweeps=30, symmetric=True, seed=0)
            res["cpu_oracle_1thread"] = {"s": time.perf_counter() - t0,
                                         "stop_reason": str(io.stop_reason),
                                         "sweeps": io.num_sweeps_run, "max_bond": max(xo.bonds),
                                         "residual_fp64": io.residuals[-1]}
    return res


def c3_diffusion_3d(with_cpu: bool, solve_levels=(4, 5), cpu_level: int = 4, full_level: int = 12):
    """C3: -div(exp(g) grad u) = 1 on a (2**L)^3 grid (3L QTT cores), g a smooth
    random rank-8 field over the three axes (two sine modes per factor, RMS 0.5,
    seed 0), k = exp(g) at the faces by scaling and squaring with sketched
    rounding (rank cap 48, error 1e-8), operator rounded at 1e-10.  The operator
    is built at the full 4096^3 size (L=12, 36 cores) and timed; AMEn
    (max_rank=64, residual_tol=1e-8, max_sweeps=30, symmetric, seed 0,
    x0 = rand(ranks=2, seed=0)) runs at the reduced sizes, and the oracle's AMEn on
    the same operator cores at ``cpu_level`` on all host cores."""
    import os
    import torch
    import ttnumpy_amd as tx
    from ttnumpy_amd import device as D
    from ttnumpy_amd import diffusion3d as T3
    kw_op = dict(k_error=1e-8, k_rank=48, op_error=1e-10, method="randomized")
    kw = dict(max_rank=64, residual_tol=1e-8, max_sweeps=30, symmetric=True, seed=0)
    res = {"field": {"modes": 2, "rank": 8, "rms": 0.5, "seed": 0}, "operator_build": kw_op,
           "solve": kw}
    A, info = T3.operator(T3.FieldSpec(n_bits=full_level), **kw_op)
    res[f"build_L{full_level}"] = {"grid": f"{2**full_level}^3", "cores": 3 * full_level,
                                   "build_s": info.build_s, "op_max_bond": max(info.bonds),
                                   "op_bonds": info.bonds, "squarings": info.squarings,
                                   "k_max_bond": max(max(v) for v in info.k_bonds.values())}
    del A
    ops = {}
    from ttnumpy_amd import autotune
    for L in sorted(set(solve_levels) | ({cpu_level} if with_cpu else set())):
        A, info = T3.operator(T3.FieldSpec(n_bits=L), **kw_op)
        ops[L] = A
        if L not in solve_levels:
            continue
        if L == min(solve_levels):     # untimed run of the smallest solve, tiling tuned on it
            res["gemm_tiling"] = autotune.tune(lambda: tx.amen(
                A, T3.rhs(L), x0=tx.rand([2] * (3 * L), ranks=2, seed=0), **kw))
        x0 = tx.rand([2] * (3 * L), ranks=2, seed=0)
        torch.cuda.synchronize()
        f0 = D.flops()
        t0 = time.perf_counter()
        x, inf = tx.amen(A, T3.rhs(L), x0=x0, return_info=True, **kw)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gf = (D.flops() - f0) / 1e9
        res[f"solve_L{L}"] = {"grid": f"{2**L}^3", "cores": 3 * L, "op_max_bond": max(info.bonds),
                              "gpu_s": wall, "stop_reason": str(inf.stop_reason),
                              "sweeps": inf.num_sweeps_run, "max_bond": max(x.bonds),
                              "residual_fp64": inf.residuals[-1], "local_solves": inf.local_solves,
                              "gflop": gf, "gflop_per_s": gf / wall,
                     "frac_fp64_peak": gf / wall / (1e3 * PEAK_FP64_TFLOPS)}
    if with_cpu:
        from threadpoolctl import threadpool_limits
        from oracle import tt_oracle as O
        L = cpu_level
        A = ops[L]
        x0 = tx.rand([2] * (3 * L), ranks=2, seed=0)
        cores = len(os.sched_getaffinity(0))
        with threadpool_limits(cores):
            t0 = time.perf_counter()
            xo, io = O.amen(O.Train(A.to_numpy()), O.Train(T3.rhs(L).to_numpy()),
                            O.Train(x0.to_numpy()), **kw)
            res[f"cpu_oracle_L{L}"] = {"s": time.perf_counter() - t0, "threads": cores,
                                       "stop_reason": str(io.stop_reason),
                                       "sweeps": io.num_sweeps_run, "max_bond": max(xo.bonds),
                                       "residual_fp64": io.residuals[-1]}
    return res


def c3_full(sweeps: int, level: int = 12):
    """C3 at its configured size: the 4096^3 operator (36 cores) built as in
    c3_diffusion_3d, then AMEn (max_rank=64, residual_tol=1e-8, symmetric,
    seed 0, x0 = rand(ranks=2, seed=0)) for at most ``sweeps`` sweeps on the
    per-kernel path.  Each sweep's FP64 residual check forms the residual
    train A x - b (bonds up to R r + 1 ~ 1.4e4) and norms it by a Q-less QR
    sweep.  The solver's own per-sweep log lines (sweep, residual, max bond,
    seconds) go to stderr as the run's progress and into the result."""
    import logging
    import torch
    import ttnumpy_amd as tx
    from ttnumpy_amd import device as D
    from ttnumpy_amd import diffusion3d as T3
    kw_op = dict(k_error=1e-8, k_rank=48, op_error=1e-10, method="randomized")
    kw = dict(max_rank=64, residual_tol=1e-8, max_sweeps=sweeps, symmetric=True, seed=0)
    lines = []

    class Keep(logging.Handler):
        def emit(self, record):
            msg = record.getMessage()
            lines.append(msg)
            print(f"# c3 full: {msg}", file=sys.stderr, flush=True)

    log = logging.getLogger("ttnumpy_amd")
    h = Keep()
    log.addHandler(h)
    old_level = log.level
    log.setLevel(logging.DEBUG)
    try:
        A, info = T3.operator(T3.FieldSpec(n_bits=level), **kw_op)
        print(f"# c3 full: operator built in {info.build_s:.1f} s, max bond {max(info.bonds)}",
              file=sys.stderr, flush=True)
        x0 = tx.rand([2] * (3 * level), ranks=2, seed=0)
        b = T3.rhs(level)
        torch.cuda.synchronize()
        f0 = D.flops()
        t0 = time.perf_counter()
        x, inf = tx.amen(A, b, x0=x0, return_info=True, **kw)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gf = (D.flops() - f0) / 1e9
        return {"grid": f"{2**level}^3", "cores": 3 * level, "op_max_bond": max(info.bonds),
                "op_build_s": info.build_s, "solve": kw, "gpu_s": wall,
                "stop_reason": str(inf.stop_reason), "sweeps": inf.num_sweeps_run,
                "residuals_fp64": inf.residuals, "max_bonds": inf.max_bonds,
                "final_bonds": x.bonds, "local_solves": inf.local_solves,
                "local_inexact": inf.local_inexact, "gflop": gf, "gflop_per_s": gf / wall,
                "frac_fp64_peak": gf / wall / (1e3 * PEAK_FP64_TFLOPS),
                "sweep_log": lines}
    finally:
        log.removeHandler(h)
        log.setLevel(old_level)


def c4_serving(n_requests: int, with_cpu: bool, cpu_per_level: int):
    """C4: the seeded Poisson-arrival stream of parametric 2D diffusion solves
    (``ttnumpy_amd.serving``) on one engine per grid level; with ``with_cpu``
    the oracle (one thread) solves ``cpu_per_level`` requests of each level
    and a FIFO single-server queue over the same arrival times, with each
    request's service time the mean sampled time of its level, gives the
    CPU's p50/p99 and solves/s."""
    import dataclasses
    from ttnumpy_amd import problems as P
    from ttnumpy_amd.serving import (ServingConfig, build_requests, fifo_latencies, request_stream,
                                     serve)
    cfg = ServingConfig(n_requests=n_requests)
    items = request_stream(cfg)
    reqs = build_requests(cfg, items)
    t0 = time.perf_counter()

    def progress(k, r):
        if r is None:
            print(f"# c4: engine tuned at {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
        elif k % 64 == 0 or k == n_requests:
            print(f"# c4: {k} done at {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    rep, _ = serve(cfg, reqs, progress=progress)
    out = {"gpu": dataclasses.asdict(rep)}
    # the same arrivals in the mixed form (FP32 Cholesky / FP32-operator MINRES
    # with FP64 refinement; environments, residual check, QR/SVD and rank rule FP64)
    cfgm = ServingConfig(n_requests=n_requests, precision="mixed")
    t0 = time.perf_counter()
    repm, _ = serve(cfgm, build_requests(cfgm, items), progress=progress)
    out["gpu_mixed"] = dataclasses.asdict(repm)
    out["mixed_ratio_vs_fp64"] = repm.solves_per_s / rep.solves_per_s
    for key in ("gpu", "gpu_mixed"):
        out[key]["frac_fp64_peak"] = out[key]["gflop_per_s"] / (1e3 * PEAK_FP64_TFLOPS)
    if with_cpu:
        from threadpoolctl import threadpool_limits
        from oracle import tt_oracle as O
        service = {}
        stops = {}
        for L in sorted(set(it.level for it in items)):
            sample = [it for it in items if it.level == L][:cpu_per_level]
            parts = P.diffusion_2d_parts(L)
            times = []
            for it in sample:
                A = O.Train(P.diffusion_2d_cores(L, it.theta, parts))
                b = O.Train(P.ones_cores([2] * (2 * L)))
                x0 = O.Train(P.random_train_cores([2] * (2 * L), 2, it.index))
                with threadpool_limits(1):
                    ts = time.perf_counter()
                    _, io = O.amen(A, b, x0, symmetric=True, residual_tol=cfg.residual_tol,
                                   max_sweeps=cfg.max_sweeps, seed=it.index)
                    times.append(time.perf_counter() - ts)
                stops.setdefault(L, []).append(str(io.stop_reason))
            service[L] = float(np.mean(times))
        finish, lat = fifo_latencies([it.offset for it in items],
                                     [service[it.level] for it in items])
        out["cpu_oracle_fifo_1thread"] = {
            "solves_per_s": n_requests / finish[-1], "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "service_s_by_level": service, "stops_by_level": stops,
            "sample": f"{cpu_per_level} oracle solves per level, FIFO queue simulated over the "
                      f"same {n_requests} arrivals"}
    return out


def cpu_baseline(budget_s: float = 8.0):
    """The oracle's AMEn (FP64 NumPy/SciPy) on the same C1 inputs, timed with
    OpenBLAS on all host cores of this process and on one thread; the faster
    of the two is the reported value (the local contractions are small, so
    threaded BLAS loses).  Also the process-parallel bound, cores x the
    one-thread rate (independent solves, one per core, perfect scaling)."""
    from threadpoolctl import threadpool_limits
    from oracle import tt_oracle as O
    from ttnumpy_amd import problems as P
    d = 2 * N_BITS
    A = O.Train(P.poisson_2d_cores(N_BITS))
    b = O.Train(P.ones_cores([2] * d))
    x0 = O.Train(P.random_train_cores([2] * d, 2, 0))
    cores = len(os.sched_getaffinity(0))

    def timed(threads):
        times = []
        with threadpool_limits(threads):
            t_end = time.perf_counter() + budget_s
            while True:
                t0 = time.perf_counter()
                O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0)
                times.append(time.perf_counter() - t0)
                if time.perf_counter() > t_end or len(times) >= 20:
                    break
        return times

    t_one, t_all = timed(1), timed(cores)
    one, many = len(t_one) / sum(t_one), len(t_all) / sum(t_all)
    best_threads, best = (1, one) if one >= many else (cores, many)
    return {"value": best, "unit": "solves/s", "cores": best_threads, "kind": "port",
            "sample": f"C1 solves, oracle FP64 NumPy/SciPy: {len(t_one)} on 1 OpenBLAS thread "
                      f"(median {np.median(t_one):.3f} s per solve), {len(t_all)} on {cores} "
                      f"threads (median {np.median(t_all):.3f} s)",
            "value_1thread": one, "value_all_threads": many, "host_cores": cores,
            "process_parallel_bound": cores * one}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--poll", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-c2", action="store_true", help="skip the C2 rounding measurement")
    ap.add_argument("--no-c5", action="store_true", help="skip the C5 high-dimensional measurement")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 3-D diffusion measurement")
    ap.add_argument("--c4", type=int, default=0, metavar="N",
                    help="instead of the C1 line, serve the C4 stream of N requests and print its "
                         "report (with the oracle FIFO estimate unless --no-cpu)")
    ap.add_argument("--c3-full-sweeps", type=int, default=0, metavar="N",
                    help="instead of the C1 line, run C3 at 4096^3 (36 cores) for at most N sweeps "
                         "and print its report")
    ap.add_argument("--c4-requests", type=int, default=512,
                    help="requests of the C4 serving leg of the full line (0 skips it)")
    ap.add_argument("--c4-cpu-per-level", type=int, default=2)
    ap.add_argument("--only", choices=("c2", "c3", "c5"), default=None,
                    help="instead of the C1 line, run one of the other measurements and print it")
    ap.add_argument("--no-single", action="store_true",
                    help="skip timing one request alone (its launches are not part of the roofline)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent solves would run one engine per GPU, without exchange)", file=sys.stderr)
        sys.exit(2)
    import torch
    from ttnumpy_amd.engine import Engine
    assert torch.cuda.is_available(), "bench.py needs the MI355X"
    if args.only is not None:
        leg = {"c2": ("c2_rounding", c2_rounding), "c3": ("c3_diffusion_3d", c3_diffusion_3d),
               "c5": ("c5_high_dim", c5_high_dim)}[args.only]
        print(json.dumps({leg[0]: leg[1](not args.no_cpu)}), flush=True)
        return
    if args.copy_calibration:
        from ttnumpy_amd import device as D
        n = args.copy_calibration * (1 << 20) // 8
        src, dst = D.empty(n), D.empty(n)
        D.fill(src, 1.0)
        for _ in range(3):
            D.copy_nd([n], D.P(src), [1], D.P(dst), [1])
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": 8 * n, "launches": 3}))
        return
    if args.c4:
        print(json.dumps({"c4_serving": c4_serving(args.c4, not args.no_cpu, args.c4_cpu_per_level)}),
              flush=True)
        return
    if args.c3_full_sweeps:
        print(json.dumps({"c3_full": c3_full(args.c3_full_sweeps)}), flush=True)
        return
    A, b, x0 = build_inputs()
    eng = Engine(args.batch, rmax=RMAX)
    # tile variant of the engine's kernels, picked on one untimed batch by
    # the engine tuner (both variants compute the same values)
    from ttnumpy_amd.autotune import tune_engine
    tiles = tune_engine(eng, requests(A, b, x0, args.batch))
    print(f"# engine tiles {tiles['tiles']} (gain {tiles['gain_vs_default']:.3f})",
          file=sys.stderr, flush=True)
    value, lat, done, total = timed_batches(eng, A, b, x0, args, "fp64")
    errors = [r.error for r in done if r.error]
    ok = [r for r in done if r.error is None]
    nsolved = len(ok)

    single_ms = None
    if not args.no_single:
        # one request alone (latency floor of a single solve on this engine)
        single = Engine(1, rmax=RMAX)
        single.run(requests(A, b, x0, 1), poll=args.poll)
        t1 = time.perf_counter()
        single.run(requests(A, b, x0, 1), poll=args.poll)
        torch.cuda.synchronize()
        single_ms = 1e3 * (time.perf_counter() - t1)

    # dominant-kernel roofline: one more batch replayed tick by tick with HIP
    # events around each kernel on the engine's stream; algorithmic flops are
    # banked per kernel by the device counters of each slot
    prof = Engine(args.batch, rmax=RMAX)
    prof.tiles = eng.tiles
    prof.profiling = True
    prof.run(requests(A, b, x0, args.batch), poll=args.poll)
    j = int(np.argmax(prof.kernel_ms))
    launches = prof.kernel_launches
    avg_ms = prof.kernel_ms[j] / launches
    flops_per_launch = prof.kernel_flops[j] / launches
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    traffic, traffic_src = measured_traffic(KERNELS[j])
    roof = {"bound": "mfma", "kernel": KERNELS[j], "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
            "traffic_source": traffic_src,
            "hbm_gbs_at_traffic": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
            "avg_launch_us": 1e3 * avg_ms, "flops_per_launch": flops_per_launch,
            "launches": launches, "kernel_ms_share": {
                k.split("::")[1]: round(float(v / prof.kernel_ms.sum()), 4)
                for k, v in zip(KERNELS, prof.kernel_ms)}}
    r0 = ok[0] if ok else None
    verify = verify_result(A, b, r0) if r0 is not None else None
    mixed = None
    if not args.no_mixed:
        # the mixed-precision engine on the same batches: FP32 Cholesky with FP64
        # refinement for the direct local solves (N >= 256), MINRES on the FP32
        # operator with FP64 defect correction for the iterative ones; the
        # environments, residual check, QR/SVD and rank rule stay FP64
        engm = Engine(args.batch, rmax=RMAX)
        engm.tiles = eng.tiles
        mval, mlat, mdone, _ = timed_batches(engm, A, b, x0, args, "mixed")
        mok = [r for r in mdone if r.error is None]
        m0 = mok[0] if mok else None
        mixed = {"value": mval, "unit": "solves/s", "ratio_vs_fp64": mval / value if value else None,
                 "p50_ms": 1e3 * float(np.percentile(mlat, 50)),
                 "p99_ms": 1e3 * float(np.percentile(mlat, 99)),
                 "solves_ok": len(mok), "errors": [r.error for r in mdone if r.error][:3],
                 "mixed_direct_solves_per_solve": engm.mixed_solves / max(len(mdone), 1),
                 "refinement_steps_per_mixed_solve": engm.mixed_refines / max(engm.mixed_solves, 1),
                 "fp64_fallbacks": engm.mixed_fallbacks}
        if m0 is not None and r0 is not None:
            from oracle import tt_oracle as O
            mv = verify_result(A, b, m0)
            xm, x64 = O.Train(m0.result.to_numpy()), O.Train(r0.result.to_numpy())
            diff = O.tt_norm(O.tt_add(xm, O.tt_scale(x64, -1.0))) / O.tt_norm(x64)
            mixed.update({"final_residual": m0.info.residuals[-1], "sweeps": m0.info.num_sweeps_run,
                          "stop_reason": str(m0.info.stop_reason), "verify": mv,
                          "relative_difference_vs_fp64_solution": float(diff),
                          "mixed_minres_iters_per_solve": engm.minres_iters / max(len(mdone), 1)})
            if not args.no_cpu:
                # distance to the oracle's FP64 solution of C1 (one BLAS thread, the
                # trajectory of the reference's own C1 run)
                from threadpoolctl import threadpool_limits
                from ttnumpy_amd import problems as P
                with threadpool_limits(1):
                    xo, _ = O.amen(O.Train(P.poisson_2d_cores(N_BITS)),
                                   O.Train(P.ones_cores([2] * (2 * N_BITS))),
                                   O.Train(x0.to_numpy()), symmetric=True, residual_tol=1e-8,
                                   max_sweeps=20, seed=0)
                for key, xx in (("mixed", xm), ("fp64", x64)):
                    mixed[f"relative_difference_vs_oracle_solution_{key}"] = float(
                        O.tt_norm(O.tt_add(xx, O.tt_scale(xo, -1.0))) / O.tt_norm(xo))
    line = {
        "metric": "AMEn solves/s + p99 latency, mixed vs FP64 on same GPU at residual 1e-8",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C1: 2D Poisson 256x256 QTT (d=16, TTM bonds<=4), b=ones, "
                               "x0=rand(ranks=2,seed=0), amen(symmetric, residual_tol=1e-8, "
                               "max_sweeps=20, seed=0), FP64",
                   "batch": args.batch, "engine": "device-resident, one slot per request",
                   "rank_capacity": RMAX},
        "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "single_request_ms": single_ms,
        "engine_tiles": {"tiles": tiles["tiles"], "variant": tiles["variant"],
                         "gain_vs_default": tiles["gain_vs_default"],
                         "kernel_ms_total": {v: k["total"] for v, k in tiles["kernel_ms"].items()},
                         "chol_kernel_ms": {v: k["eng_chol_kernel"]
                                            for v, k in tiles["kernel_ms"].items()}},
        "solves_ok": nsolved, "errors": errors[:3],
        "final_residual": r0.info.residuals[-1] if r0 else None,
        "sweeps": r0.info.num_sweeps_run if r0 else None,
        "stop_reason": str(r0.info.stop_reason) if r0 else None,
        "max_bond": max(r0.result.bonds) if r0 else None,
        "minres_iters_per_solve": prof.minres_iters / max(args.batch, 1),
        "lu_solves_per_solve": prof.lu_solves / max(args.batch, 1),
        "gflop_per_solve": float(prof.kernel_flops.sum()) / max(args.batch, 1) / 1e9,
        # every engine kernel's algorithmic flops at the timed solve rate, against
        # the FP64 MFMA peak (the roofline object is the dominant kernel alone)
        "engine_tflops": float(prof.kernel_flops.sum()) / max(args.batch, 1) * value / 1e12,
        "engine_frac_fp64_peak": float(prof.kernel_flops.sum()) / max(args.batch, 1) * value
                                 / (1e12 * PEAK_FP64_TFLOPS),
        "stage_ms_per_solve": {nm: round(float(v) / max(args.batch, 1), 3) for nm, v in zip(
            ("rhs", "matrix", "lu", "minres", "exact_res", "post", "check", "env", "post.residual_core",
             "post.svd", "post.rank", "post.enrich_qr", "post.zshift", "lu.panels", "lu.in_panel",
             "lu.block_rows", "lu.trailing", "lu.backsub"), prof.stage_ms)},
        "roofline": roof,
        "verify": verify,
        "mixed": mixed,
    }
    if not args.no_c2:
        try:
            line["c2_rounding"] = c2_rounding(not args.no_cpu)
        except Exception as exc:  # the C1 line is reported whatever happens here
            line["c2_rounding"] = {"error": repr(exc)}
        print("# c2 done", file=sys.stderr, flush=True)
    if not args.no_c5:
        try:
            line["c5_high_dim"] = c5_high_dim(not args.no_cpu)
            if mixed is not None:
                mixed["c5_engine_one_slot"] = line["c5_high_dim"].get("engine_mixed")
                mixed["c5_engine_one_slot_bf16"] = line["c5_high_dim"].get("engine_bf16")
        except Exception as exc:  # the C1 line is reported whatever happens here
            line["c5_high_dim"] = {"error": repr(exc)}
        print("# c5 done", file=sys.stderr, flush=True)
    if not args.no_c3:
        try:
            line["c3_diffusion_3d"] = c3_diffusion_3d(not args.no_cpu)
        except Exception as exc:  # the C1 line is reported whatever happens here
            line["c3_diffusion_3d"] = {"error": repr(exc)}
        print("# c3 done", file=sys.stderr, flush=True)
    if args.c4_requests:
        try:
            line["c4_serving"] = c4_serving(args.c4_requests, not args.no_cpu,
                                            args.c4_cpu_per_level)
        except Exception as exc:  # the C1 line is reported whatever happens here
            line["c4_serving"] = {"error": repr(exc)}
        print("# c4 done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
