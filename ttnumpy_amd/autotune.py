"""Autotuners: the GEMM tiling of the per-kernel path, the engine's tile variants.

``ttx_gemm_f64`` runs a contraction on 64 x 64 output tiles, or on 32 x 32 tiles
(four times the waves, a quarter of the MFMA chain each, identical values) when
its 64 x 64 grid has fewer tiles than a process-wide limit.  Which is faster
depends on the shape and on the GPU, so the limit is chosen from timings:

1. ``recording()`` collects the shapes of the GEMMs a workload issues (the
   local matvecs issued from C included) with their call counts;
2. ``time_shapes`` times every distinct shape under each tiling with HIP
   events on the current stream;
3. ``choose_limit`` picks the limit that minimises the call-weighted sum of
   the timings, and ``tune`` sets it (``ttx_gemm_set_small_tile_limit``).

    python -m ttnumpy_amd.autotune [--workload c5|c3|c2] [--reps 20] [--save FILE]

``tune_engine`` picks the batched engine's tile variant (128- or 64-row
tiles below the Cholesky factorization's diagonal blocks; identical values)
from the engine's own kernel times on a sample of the requests it will serve.
"""

from __future__ import annotations

import argparse
import contextlib
import json
import time
from typing import Callable, Dict, Optional, Tuple

import torch

from . import _lib
from . import device as D
from .device import P

Shape = Tuple[int, int, int, int, bool, bool]   # batch, M, N, K, A m-fastest, B n-fastest

ALL_64 = 0            # limit under which no grid falls: every contraction on 64 x 64
ALL_32 = 1 << 62      # limit above every grid: every contraction on 32 x 32


@contextlib.contextmanager
def recording():
    """Collect ``{shape: calls}`` of the FP64 GEMMs issued inside the block."""
    rec: Dict[Shape, int] = {}
    prev = D._GEMM_RECORD[0]
    D._GEMM_RECORD[0] = rec
    try:
        yield rec
    finally:
        D._GEMM_RECORD[0] = prev


def tiles64(shape: Shape) -> int:
    batch, M, N = shape[:3]
    return batch * ((M + 63) // 64) * ((N + 63) // 64)


def get_limit() -> int:
    return int(D.lib().ttx_gemm_small_tile_limit())


def set_limit(tiles: int) -> None:
    _lib.check(D.lib().ttx_gemm_set_small_tile_limit(int(tiles)), "ttx_gemm_set_small_tile_limit")


def time_shapes(shapes, reps: int = 20, warmup: int = 3) -> Dict[Shape, Tuple[float, float]]:
    """``{shape: (ms on 32 x 32, ms on 64 x 64)}``, the mean of ``reps`` launches each."""
    old = get_limit()
    out = {}
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    try:
        for shape in shapes:
            batch, M, N, K, a_m_fast, b_n_fast = shape
            A = D.empty(batch * M * K).normal_()
            B = D.empty(batch * K * N).normal_()
            C = D.empty(batch * M * N)
            a_ms, a_ks = (1, M) if a_m_fast else (K, 1)
            b_ks, b_ns = (N, 1) if b_n_fast else (1, K)
            ms = []
            for limit in (ALL_32, ALL_64):
                set_limit(limit)
                for _ in range(warmup):
                    D.lib().ttx_gemm_f64(D.stream(), batch, M, N, K, 1.0, P(A), a_ms, a_ks, M * K,
                                         P(B), b_ks, b_ns, K * N, 0.0, P(C), N, 1, M * N)
                start.record()
                for _ in range(reps):
                    D.lib().ttx_gemm_f64(D.stream(), batch, M, N, K, 1.0, P(A), a_ms, a_ks, M * K,
                                         P(B), b_ks, b_ns, K * N, 0.0, P(C), N, 1, M * N)
                stop.record()
                stop.synchronize()
                ms.append(start.elapsed_time(stop) / reps)
            out[shape] = (ms[0], ms[1])
    finally:
        set_limit(old)
    return out


def choose_limit(times: Dict[Shape, Tuple[float, float]], calls: Dict[Shape, int]) -> Tuple[int, float]:
    """The limit (in 64 x 64 tiles) with the least call-weighted time, and that time (ms)."""
    cands = sorted({0} | {tiles64(s) + 1 for s in times})
    best, best_ms = None, None
    for lim in cands:
        tot = sum(calls.get(s, 1) * (t32 if tiles64(s) < lim else t64) for s, (t32, t64) in times.items())
        if best_ms is None or tot < best_ms - 1e-12:
            best, best_ms = lim, tot
    return best, best_ms


def predicted_ms(times, calls, limit: int) -> float:
    return sum(calls.get(s, 1) * (t32 if tiles64(s) < limit else t64) for s, (t32, t64) in times.items())


def tune(workload: Callable[[], object], reps: int = 20, apply: bool = True) -> dict:
    """Run ``workload`` once while recording its GEMMs, time their shapes under
    both tilings and (with ``apply``) set the best limit.  Returns a report."""
    with recording() as rec:
        workload()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    times = time_shapes(rec.keys(), reps=reps)
    limit, best = choose_limit(times, rec)
    default = get_limit()
    report = {"limit": limit, "previous_limit": default, "distinct_shapes": len(rec),
              "calls": int(sum(rec.values())), "tuning_s": time.perf_counter() - t0,
              "predicted_ms": {"tuned": best, "previous": predicted_ms(times, rec, default),
                               "all_64x64": predicted_ms(times, rec, ALL_64),
                               "all_32x32": predicted_ms(times, rec, ALL_32)}}
    if apply:
        set_limit(limit)
    return report


# ------------------------------------------------------------------ engine tiles
# The engine's kernels take their tile variants from each slot's `tiles`
# bits; the variants give identical values, so they can be told apart by
# speed alone, on the requests an engine is about to serve.
ENGINE_TILE_VARIANTS = {
    0: "Cholesky rows below each diagonal block in tiles of 128 rows x 64 columns",
    1: "Cholesky rows below each diagonal block in tiles of 64 rows x 64 columns",
}


def tune_engine(engine, sample, poll: int = 64, ticks: Optional[int] = None,
                apply: bool = True) -> dict:
    """Pick ``engine``'s tile variant from its own kernels' times on ``sample``.

    For each variant a scratch engine of the same slot count and rank
    capacity serves copies of the ``sample`` requests (up to its slots) with
    HIP events around each of the seven kernels of every tick: to completion,
    or for ``ticks`` ticks (a bound on the tuning time for long solves).  The
    variants compute the same values, so both do the same work, and the
    variant with the least total kernel time wins.  Returns a report
    (per-variant kernel milliseconds, the choice, the gain over variant 0)
    and, with ``apply``, sets ``engine.tiles``."""
    import dataclasses
    import numpy as np
    from .engine import KERNEL_NAMES, Engine
    t0 = time.perf_counter()
    reqs = list(sample)[:engine.nslots]
    res = {}
    for v in ENGINE_TILE_VARIANTS:
        eng = Engine(engine.nslots, rmax=engine.rmax, chunk=engine.chunk,
                     hbm_reserve_bytes=engine.reserve, minres_slice=engine.minres_slice,
                     graphs=False, max_rank_capacity=engine.max_rank_capacity)
        eng.tiles = v
        eng.profiling = True
        copies = [dataclasses.replace(r, result=None, info=None, error=None, admitted=0.0,
                                      finished=0.0, arrival=0.0) for r in reqs]
        if ticks is None:
            eng.run(copies, poll=poll)
        else:  # finished slots idle through the remaining ticks
            for i, r in enumerate(copies):
                eng.admit(i, r, push=False)
            eng._push_all()
            eng.tick(ticks)
        torch.cuda.synchronize()
        res[v] = {k: round(float(x), 3) for k, x in zip(KERNEL_NAMES, eng.kernel_ms)}
        res[v]["total"] = round(float(np.sum(eng.kernel_ms)), 3)
        del eng
    best = min(res, key=lambda v: res[v]["total"])
    report = {"tiles": best, "variant": ENGINE_TILE_VARIANTS[best], "requests": len(reqs),
              "ticks": ticks,
              "kernel_ms": res, "gain_vs_default": res[0]["total"] / res[best]["total"],
              "tuning_s": time.perf_counter() - t0}
    if apply:
        engine.tiles = best
    return report


def load(path: str) -> int:
    """Set the limit saved by ``--save``; returns it."""
    with open(path) as fh:
        limit = int(json.load(fh)["limit"])
    set_limit(limit)
    return limit


def _workload(name: str) -> Callable[[], object]:
    import ttnumpy_amd as tx
    from . import problems as Pb
    if name == "c5":
        A = tx.TensorTrain(Pb.shifted_laplacian_sum_cores(64, 8, 1.0 / 64))
        b = tx.TensorTrain(Pb.random_train_cores([8] * 64, 16, 3))
        return lambda: tx.amen(A, b, max_rank=128, residual_tol=1e-8, max_sweeps=30,
                               symmetric=True, seed=0)
    if name == "c3":
        from . import diffusion3d as T3
        A, _ = T3.operator(T3.FieldSpec(n_bits=4), k_error=1e-8, k_rank=48, op_error=1e-10,
                           method="randomized")
        return lambda: tx.amen(A, T3.rhs(4), x0=tx.rand([2] * 12, ranks=2, seed=0), max_rank=64,
                               residual_tol=1e-8, max_sweeps=30, symmetric=True, seed=0)
    if name == "c2":
        x1 = tx.TensorTrain(Pb.random_train_cores([4] * 40, 64, 1, normal=True))
        x2 = tx.TensorTrain(Pb.random_train_cores([4] * 40, 64, 2, normal=True))
        return lambda: (x1 + x2).truncate(error=1e-10, max_rank=96)
    raise ValueError(f"unknown workload {name!r}")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=("c5", "c3", "c2"), default="c5")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--save", default=None, help="write the chosen limit to this JSON file")
    a = ap.parse_args(argv)
    report = tune(_workload(a.workload), reps=a.reps)
    report["workload"] = a.workload
    print(json.dumps(report), flush=True)
    if a.save:
        with open(a.save, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    main()
