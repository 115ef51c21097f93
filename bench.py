"""Benchmark: AMEn solves/s and latency on the 2D Poisson 256x256 QTT system, one MI355X.

Workload (one "step" = one full solve, inputs resident in HBM before the
clock starts): A = -Laplace on a 256 x 256 grid (Dirichlet, h = 1/257) as
the Kronecker-sum QTT operator of 16 cores with 2x2 modes and TTM bonds <= 4,
b = all-ones rank-1 train, x0 = rand([2]*16, ranks=2, seed=0);
amen(symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0) in FP64.

Prints one JSON line (rank 0).  ``--gpus N`` with N > 1 exits: this build
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
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def build_inputs():
    import ttnumpy_amd as tx
    from ttnumpy_amd import problems as P
    d = 2 * N_BITS
    A = tx.TensorTrain(P.poisson_2d_cores(N_BITS))
    b = tx.TensorTrain(P.ones_cores([2] * d))
    x0 = tx.rand([2] * d, ranks=2, seed=0)
    return A, b, x0


def solve(A, b, x0):
    import ttnumpy_amd as tx
    return tx.amen(A, b, x0=x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0,
                   return_info=True)


def kernel_roofline(A, b, x0, kernel: str = "gemm"):
    """Per-launch timing of the dominant kernel over one extra solve, with HIP events.

    Events are recorded on the stream the kernels are launched on (torch's
    current stream, which ttx launches onto).  Algorithmic flops per launch
    are 2*M*N*K*batch for the GEMM.
    """
    import torch
    from ttnumpy_amd import device as D
    rec = []
    orig = D.gemm

    def traced(batch, M, N, K, *args):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        orig(batch, M, N, K, *args)
        e.record()
        rec.append((2.0 * batch * M * N * K, s, e))

    D.gemm = traced
    try:
        solve(A, b, x0)
    finally:
        D.gemm = orig
    torch.cuda.synchronize()
    flops = sum(r[0] for r in rec)
    ms = sum(r[1].elapsed_time(r[2]) for r in rec)
    launches = len(rec)
    avg_us = 1e3 * ms / max(launches, 1)
    achieved = (flops / launches) / (avg_us * 1e-6) / 1e12 if launches else 0.0
    return {
        "bound": "mfma", "kernel": "ttx::gemm_f64_kernel", "achieved": achieved,
        "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
        "traffic": None, "launches_per_solve": launches, "avg_launch_us": avg_us,
        "flops_per_launch": flops / max(launches, 1),
    }


def cpu_baseline(budget_s: float = 12.0):
    """The oracle's AMEn on the same C1 inputs on the host, 1 BLAS thread."""
    from threadpoolctl import threadpool_limits
    from oracle import tt_oracle as O
    from ttnumpy_amd import problems as P
    d = 2 * N_BITS
    A = O.Train(P.poisson_2d_cores(N_BITS))
    b = O.Train(P.ones_cores([2] * d))
    x0 = O.Train(P.random_train_cores([2] * d, 2, 0))
    times = []
    with threadpool_limits(1):
        t_end = time.perf_counter() + budget_s
        while True:
            t0 = time.perf_counter()
            O.amen(A, b, x0, symmetric=True, residual_tol=1e-8, max_sweeps=20, seed=0)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() > t_end or len(times) >= 20:
                break
    return {"value": len(times) / sum(times), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{len(times)} C1 solves (oracle FP64 NumPy/SciPy, OpenBLAS 1 thread), "
                      f"median {np.median(times):.3f}s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is documented, not built)", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs the MI355X"
    A, b, x0 = build_inputs()
    for _ in range(args.warmup):
        solve(A, b, x0)
    torch.cuda.synchronize()
    lat = []
    info = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        x, info = solve(A, b, x0)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - s)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    value = args.steps / total
    roof = kernel_roofline(A, b, x0)
    line = {
        "metric": "AMEn solves/s + p99 latency, mixed vs FP64 on same GPU at residual 1e-8",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C1: 2D Poisson 256x256 QTT (d=16, TTM bonds<=4), b=ones, "
                               "x0=rand(ranks=2,seed=0), AMEn residual_tol=1e-8 FP64",
                   "precision": "fp64-plain"},
        "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "final_residual": info.residuals[-1], "sweeps": info.num_sweeps_run,
        "stop_reason": str(info.stop_reason), "max_bond": max(x.bonds),
        "roofline": roof,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
