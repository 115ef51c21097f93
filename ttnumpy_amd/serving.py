"""Serving loop for parametric 2-D diffusion solves on the batched engine.

    python -m ttnumpy_amd.serving --requests 512 [--rate 50] [--precision mixed]

A stream of requests arrives with exponential inter-arrival times (Poisson
arrivals at ``rate`` per second).  Request ``i`` solves
``-div(k grad u) = 1`` on a ``2**L x 2**L`` grid with
``k = 1 + theta sin(pi x) sin(pi y)``, ``L`` uniform in ``levels`` and
``theta`` uniform in ``[0, theta_max)`` (:func:`ttnumpy_amd.problems.diffusion_2d_cores`),
by AMEn (``symmetric=True``, ``residual_tol``, ``max_sweeps``) from
``x0 = rand([2]*2L, ranks=2, seed=i)``.

The engine admits a request once it has arrived, a slot is free and free HBM
covers the slot's arena.  Each poll is one launch in which every slot runs the
steps of its own sweep at its own pace for about ``pace_us`` (the fused kernel
of all seven stages, called pass after pass; ``pace_us=0``: ``poll`` ticks in
lockstep, one launch per tick); the results are the bits the tick kernels give.
With ``per_level`` each grid level has its own engine on its own stream.
Latency is measured from arrival to result.  Operators and right-hand sides are built
and copied to the device before the clock starts, as a client would hand
them over.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import device as D
from . import problems as P
from .engine import Engine, EnginePool, Request
from .tensor_train import TensorTrain, rand


@dataclass
class ServingConfig:
    n_requests: int = 512
    rate: float = 50.0            # arrivals per second
    levels: Sequence[int] = (8, 9, 10, 11, 12)
    theta_max: float = 0.9
    seed: int = 0
    residual_tol: float = 1e-8
    max_sweeps: int = 20
    nslots: int = 512             # per engine
    rmax: int = 64
    poll: int = 16
    # one engine (own stream) per grid level: with lockstep ticks it keeps small
    # requests from waiting on large ones; paced polls need no such split, and
    # one engine whose slots are all resident at once serves more
    per_level: bool = False
    minres_slice: int = 128       # MINRES iterations per tick
    precision: str = "fp64"       # "mixed": FP32 Cholesky with FP64 refinement for direct solves
    # before the clock starts, each engine's tile variant is picked from its
    # kernels' times over `tune_ticks` ticks of up to `tune_sample` requests
    # of its own level (0: variant 0)
    tune_sample: int = 8
    tune_ticks: int = 192
    # one kernel per tick in which each slot runs the seven stages without
    # waiting for the other slots between them (Engine(fused=True), same
    # bits): the slots of a level hold requests at different sites and sweeps
    fused: bool = True
    # with fused, pace_us > 0: each poll is one launch in which every slot runs
    # ticks at its own pace for about pace_us microseconds (Engine.pace_us;
    # same bits): no slot waits for a slower one at tick boundaries either
    pace_us: float = 100000.0


@dataclass
class ServingItem:
    index: int
    level: int
    theta: float
    offset: float                 # arrival time after the start, seconds


def request_stream(cfg: ServingConfig) -> List[ServingItem]:
    """The seeded arrivals, levels and coefficients."""
    rng = np.random.default_rng(cfg.seed)
    gaps = rng.exponential(1.0 / cfg.rate, size=cfg.n_requests)
    levels = rng.choice(np.asarray(cfg.levels), size=cfg.n_requests)
    thetas = rng.uniform(0.0, cfg.theta_max, size=cfg.n_requests)
    t = np.cumsum(gaps) - gaps[0]
    return [ServingItem(i, int(levels[i]), float(thetas[i]), float(t[i]))
            for i in range(cfg.n_requests)]


def build_requests(cfg: ServingConfig, items: Sequence[ServingItem]) -> List[Request]:
    """Device-resident operators, right-hand sides and initial guesses."""
    parts: Dict[int, tuple] = {}
    rhs: Dict[int, TensorTrain] = {}
    out = []
    for it in items:
        L = it.level
        if L not in parts:
            parts[L] = P.diffusion_2d_parts(L)
            rhs[L] = TensorTrain(P.ones_cores([2] * (2 * L)))
        A = TensorTrain(P.diffusion_2d_cores(L, it.theta, parts[L]))
        x0 = rand([2] * (2 * L), ranks=2, seed=it.index)
        out.append(Request(A=A, b=rhs[L], x0=x0, symmetric=True, residual_tol=cfg.residual_tol,
                           max_sweeps=cfg.max_sweeps, seed=it.index, tag=it,
                           precision=cfg.precision))
    return out


@dataclass
class ServingReport:
    solves: int
    errors: int
    wall_s: float
    solves_per_s: float
    p50_ms: float
    p99_ms: float
    gflop: float
    gflop_per_s: float
    converged: int
    stagnated: int
    max_sweeps: int
    worst_residual: float
    by_level: Dict[int, dict] = field(default_factory=dict)
    # per engine (keyed by grid level, or "all"): the tile variant chosen on
    # the warm-up sample, its measured kernel-time gain over variant 0, tuning seconds
    tiles: Dict[object, dict] = field(default_factory=dict)
    fused: bool = True            # ticks as one fused kernel (else seven kernels)
    pace_us: float = 0.0          # paced polls of about this many microseconds (0: ticks)
    # slots holding a request, averaged over the run (sum of admission-to-result
    # times / wall clock), against the slots the engines have
    mean_busy_slots: float = 0.0
    slots: int = 0


def serve(cfg: ServingConfig, requests: Optional[List[Request]] = None,
          engine: Optional[Engine] = None, progress=None) -> tuple:
    """Run the stream; returns ``(report, requests)``.  ``progress(n_done, req)``
    is called after every finished request (and as ``progress(0, None)`` after
    each engine's tuning)."""
    items = request_stream(cfg)
    if requests is None:
        requests = build_requests(cfg, items)
    levels = sorted(set(int(x) for x in cfg.levels))
    if engine is not None:
        engines = [engine]
        pool = None
    elif cfg.per_level:
        engines = [Engine(cfg.nslots, rmax=cfg.rmax, minres_slice=cfg.minres_slice, fused=cfg.fused)
                   for _ in levels]
        pool = EnginePool(engines, lambda r: levels.index(r.tag.level))
    else:
        engines = [Engine(cfg.nslots, rmax=cfg.rmax, minres_slice=cfg.minres_slice, fused=cfg.fused)]
        pool = None
    for e in engines:
        e.pace_us = cfg.pace_us if e.fused else 0.0
    tiles: Dict[object, dict] = {}
    if cfg.tune_sample > 0:
        from .autotune import tune_engine
        keys = levels if pool is not None else ["all"]
        for key, eng in zip(keys, engines):
            sample = [r for r, it in zip(requests, items) if key == "all" or it.level == key]
            if sample:
                t = tune_engine(eng, sample[:cfg.tune_sample], ticks=cfg.tune_ticks)
                if progress is not None:
                    progress(0, None)
                tiles[key] = {"tiles": t["tiles"], "variant": t["variant"],
                              "gain_vs_default": t["gain_vs_default"], "sample": t["requests"],
                              "ticks": t["ticks"],
                              "kernel_ms_total": {v: k["total"] for v, k in t["kernel_ms"].items()},
                              "tuning_s": t["tuning_s"]}
    torch.cuda.synchronize()
    flops0 = sum(float(e.kernel_flops.sum()) for e in engines)
    t0 = time.perf_counter()
    for r, it in zip(requests, items):
        r.arrival = t0 + it.offset
        r.result = r.info = r.error = None
    count = [0]

    def on_done(r):
        count[0] += 1
        if progress is not None:
            progress(count[0], r)

    queue = sorted(requests, key=lambda r: r.arrival)
    if pool is not None:
        done = pool.run(queue, poll=cfg.poll, clock=time.perf_counter, on_done=on_done)
    else:
        done = engines[0].run(queue, poll=cfg.poll, clock=time.perf_counter, on_done=on_done)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([r.finished - r.arrival for r in done])
    ok = [r for r in done if r.error is None]
    stops = [str(r.info.stop_reason) for r in ok]
    by_level: Dict[int, dict] = {}
    for r in ok:
        L = r.tag.level
        e = by_level.setdefault(L, {"n": 0, "lat_ms": [], "sweeps": [], "max_bond": 0})
        e["n"] += 1
        e["lat_ms"].append(1e3 * (r.finished - r.arrival))
        e["sweeps"].append(r.info.num_sweeps_run)
        e["max_bond"] = max(e["max_bond"], max(r.result.bonds))
    for L, e in by_level.items():
        e["p50_ms"] = float(np.percentile(e.pop("lat_ms"), 50))
        e["mean_sweeps"] = float(np.mean(e.pop("sweeps")))
    gflop = (sum(float(e.kernel_flops.sum()) for e in engines) - flops0) / 1e9
    rep = ServingReport(
        solves=len(ok), errors=len(done) - len(ok), wall_s=wall, solves_per_s=len(ok) / wall,
        p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
        gflop=gflop, gflop_per_s=gflop / wall,
        converged=stops.count("converged"), stagnated=stops.count("stagnated"),
        max_sweeps=stops.count("max_sweeps"),
        worst_residual=max((r.info.residuals[-1] for r in ok), default=float("nan")),
        by_level=dict(sorted(by_level.items())), tiles=tiles, fused=all(e.fused for e in engines),
        pace_us=max(e.pace_us for e in engines),
        mean_busy_slots=float(sum(r.finished - r.admitted for r in done)) / wall,
        slots=sum(e.nslots for e in engines))
    return rep, done


def fifo_latencies(arrivals: Sequence[float], service: Sequence[float]) -> tuple:
    """One server taking requests in arrival order: ``(finish times, latencies)``
    for arrival offsets (non-decreasing) and per-request service times."""
    free = 0.0
    finish, lat = [], []
    for a, s in zip(arrivals, service):
        start = max(free, a)
        free = start + s
        finish.append(free)
        lat.append(free - a)
    return finish, lat


def main(argv=None) -> None:
    """Serve a seeded stream and print progress lines and the report as JSON."""
    import argparse
    import dataclasses
    import json
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slots", type=int, default=512, help="slots per engine")
    ap.add_argument("--per-level", action="store_true",
                    help="one engine per grid level instead of one for all")
    ap.add_argument("--precision", choices=("fp64", "mixed", "bf16"), default="fp64")
    ap.add_argument("--tune-sample", type=int, default=32,
                    help="warm-up requests per engine for the tile tuner (0: no tuning)")
    ap.add_argument("--tick-kernels", action="store_true",
                    help="seven kernels per tick instead of the fused one")
    ap.add_argument("--pace-ms", type=float, default=100.0,
                    help="paced polls of this many milliseconds (0: polls of ticks)")
    a = ap.parse_args(argv)
    cfg = ServingConfig(n_requests=a.requests, rate=a.rate, seed=a.seed, nslots=a.slots,
                        per_level=a.per_level, precision=a.precision,
                        tune_sample=a.tune_sample, fused=not a.tick_kernels,
                        pace_us=1e3 * a.pace_ms)
    t = time.perf_counter()
    reqs = build_requests(cfg, request_stream(cfg))
    print(f"built {len(reqs)} requests in {time.perf_counter() - t:.1f} s", flush=True)
    t0 = time.perf_counter()

    def progress(k, r):
        if r is None:
            print(f"tuned an engine at {time.perf_counter() - t0:.1f} s", flush=True)
        elif k % 32 == 0 or k == a.requests:
            print(f"{k} done at {time.perf_counter() - t0:.1f} s (last: L={r.tag.level}, "
                  f"latency {r.finished - r.arrival:.2f} s)", flush=True)

    rep, done = serve(cfg, reqs, progress=progress)
    out = dataclasses.asdict(rep)
    out["config"] = {"requests": a.requests, "rate_per_s": a.rate, "seed": a.seed,
                     "levels": list(cfg.levels), "theta_max": cfg.theta_max,
                     "slots_per_engine": a.slots, "per_level": cfg.per_level,
                     "precision": a.precision, "residual_tol": cfg.residual_tol,
                     "max_sweeps": cfg.max_sweeps, "tune_sample": cfg.tune_sample,
                     "fused": cfg.fused, "pace_us": cfg.pace_us}
    errs = [r.error for r in done if r.error]
    if errs:
        out["first_errors"] = errs[:5]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
