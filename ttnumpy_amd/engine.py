"""Batched, device-resident AMEn with continuous batching.

Each request occupies a *slot*: a descriptor (``include/ttx_engine.h``) plus one
arena in HBM holding its solution and residual trains at a fixed rank
capacity, its environments and its scratch.  ``ttx_engine_tick`` advances
every active slot by one site step (environments, local solve, truncation and
enrichment, or the per-sweep residual check); rank decisions and the stopping
test are taken on the device.  The host replays ticks, polls the slot status
every few ticks, retires finished requests and admits queued ones into free
slots when HBM allows -- requests at different sites, sweeps and ranks share
every launch.

The arithmetic and its order follow ``ttnumpy/solvers/amen.py:395-725``; the
result of one slot does not depend on the other slots (one workgroup per slot,
no cross-slot reductions), so a request gives the same bits alone and in a mix.
"""

from __future__ import annotations

import ctypes
import time
import warnings
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from . import device as D
from .solvers.common import SolverInfo, StopReason, _SolverSettings
from .tensor_train import TensorTrain

MAXD = 64
MAXSWEEPS = 64
NKERNELS = 7  # per tick: env, solve (rhs + MINRES), chol32, chol, lu, post, check
# stage masks of a tick (bit i launches kernel i)
STAGES_ALL, STAGE_CHOL32, STAGE_LU, FUSED = 0x7F, 0x04, 0x10, 0x80
_SSTAGE_LU, _SSTAGE_GMRES = 2, 6  # slot.sstage of a solve waiting on the LU kernel
KERNEL_NAMES = ("eng_env_kernel", "eng_solve_kernel", "eng_chol32_kernel", "eng_chol_kernel",
                "eng_lu_kernel", "eng_post_kernel", "eng_check_kernel")
SCR_COUNT = 19

PH_XORTH, PH_ZORTH, PH_DONE = 1, 2, 14
SLOT_RUNNING, SLOT_DONE, SLOT_EMPTY = 0, 1, -100
_STATUS_TEXT = {-1: "rank capacity exceeded", -2: "NaN or infinite entries",
                -3: "singular local matrix"}

_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_vp = ctypes.c_void_p


class Slot(ctypes.Structure):
    """Host mirror of ``ttx_slot`` (every field 8 bytes wide)."""

    _fields_ = (
        [(nm, _i64) for nm in ("d", "kick", "max_rank", "direct_size", "iterative", "residual_mode",
                               "max_sweeps", "track", "rmax", "nmax", "ramax", "rbmax", "bneg",
                               "allow_fallback", "chunk", "chol", "mixed", "minres_slice")]
        + [("n", _i64 * MAXD), ("ra", _i64 * (MAXD + 1)), ("rb", _i64 * (MAXD + 1))]
        + [(nm, _f64) for nm in ("residual_tol", "trunc_tol", "local_rtol", "damp", "error_threshold",
                                 "stagnation", "bnorm", "pad1")]
        + [(nm, _i64) for nm in ("cap_core", "cap_zcore", "cap_loc", "direct_cap", "lmax", "scr_size",
                                 "env_size", "tiles")]
        + [("cap_env", _i64 * 5), ("pad3", _i64 * 3), ("env_off", _i64 * 10), ("pad4", _i64 * 6),
           ("scr_off", _i64 * 32)]
        + [("A", _vp * MAXD), ("b", _vp * MAXD), ("x", _vp), ("z", _vp), ("env", _vp), ("scr", _vp)]
        + [("rx", _i64 * (MAXD + 1)), ("rz", _i64 * (MAXD + 1))]
        + [(nm, _i64) for nm in ("phase", "k", "sweep", "status", "stop_reason", "local_solves",
                                 "local_inexact", "local_fallbacks", "rank_decisions", "rank_capped",
                                 "red_last", "svd_noconv", "ticks", "minres_iters", "lu_solves",
                                 "chol_fallbacks", "mixed_solves", "mixed_refines",
                                 "mixed_fallbacks", "sstage", "chk_core", "chk_cdim")]
        + [("flops", _f64 * NKERNELS)]
        + [("residuals", _f64 * MAXSWEEPS), ("max_bonds", _i64 * MAXSWEEPS)]
        + [("clk", _i64 * 24)]
    )


def _check_layout(lib) -> None:
    lib.ttx_engine_slot_bytes.restype = _i64
    lib.ttx_engine_field_offset.restype = _i64
    lib.ttx_engine_field_offset.argtypes = [ctypes.c_int]
    lib.ttx_engine_plan.argtypes = [ctypes.c_void_p]
    lib.ttx_engine_plan.restype = ctypes.c_int
    lib.ttx_engine_tick.argtypes = [_vp, _vp, _i64, _i64, ctypes.c_uint32]
    lib.ttx_engine_tick.restype = ctypes.c_int

    lib.ttx_engine_profile.argtypes = [_vp, _vp, _i64, _i64, ctypes.POINTER(_f64)]
    lib.ttx_engine_profile.restype = ctypes.c_int
    if lib.ttx_engine_slot_bytes() != ctypes.sizeof(Slot):
        raise RuntimeError("ttx_slot layout mismatch between libttx and the host mirror")
    for i, name in enumerate(("n", "residual_tol", "cap_core", "A", "rx", "residuals", "max_bonds")):
        if lib.ttx_engine_field_offset(i) != getattr(Slot, name).offset:
            raise RuntimeError(f"ttx_slot field {name} offset mismatch")


@dataclass
class Request:
    """One AMEn solve: operator A, right-hand side b, initial guess x0 and options."""

    A: TensorTrain
    b: TensorTrain
    x0: Optional[TensorTrain] = None
    kickrank: int = 4
    error_threshold: float = 1e-7
    max_rank: Optional[int] = None
    max_sweeps: int = 1
    residual_tol: Optional[float] = None
    residual_damp: float = 2.0
    stagnation_factor: object = "auto"
    local_solver: str = "auto"
    direct_solve_size: int = 1000
    local_rtol: Optional[float] = None
    allow_dense_fallback: bool = False
    symmetric: bool = False
    seed: Optional[int] = None
    # "mixed": direct Cholesky solves (N >= 256) factor in FP32 with FP64 refinement, and
    # MINRES solves run on the FP32 operator with FP64 defect correction; "bf16": the
    # same, with the MINRES operator's products on BF16 operands (FP32 accumulation)
    precision: str = "fp64"
    # rank capacity of the request's slot (None: the engine's rmax); raised by
    # the engine when a bond outgrows it
    rank_capacity: Optional[int] = None
    tag: object = None
    arrival: float = 0.0
    # filled by the engine
    result: Optional[TensorTrain] = None
    info: Optional[SolverInfo] = None
    error: Optional[str] = None
    admitted: float = 0.0
    finished: float = 0.0


class Engine:
    """A fixed set of slots on one GPU plus the continuous-batching loop."""

    def __init__(self, nslots: int, rmax: int, chunk: int = 8, hbm_reserve_bytes: int = 2 << 30,
                 minres_slice: int = 128, graphs: bool = True, max_rank_capacity: int = 1024,
                 fused: bool = False):
        self.lib = _lib.load()
        # A request whose bond outgrows its slot's rank capacity is planned
        # again with twice the capacity (up to max_rank_capacity, and never
        # above its max_rank) and served again from its x0: the slot arena
        # grows with the plan, and the solve is deterministic, so the result
        # is the one a slot of the larger capacity gives from the start.  The
        # reference's ranks are bounded only by max_rank (amen.py:441-451).
        self.max_rank_capacity = int(max_rank_capacity)
        self.regrown = 0
        # tile variants of the engine's kernels (the slot's `tiles` bits; equal
        # values, different speed), chosen by ttnumpy_amd.autotune.tune_engine
        self.tiles = 0
        # graphs: each poll's ticks replay one recorded HIP graph instead of
        # seven launches per tick (same kernels, same order, same results)
        self.graphs = bool(graphs)
        self._graphs = {}
        # launch the FP32 and LU kernels only in polls where a resident slot
        # can use them (see _stages); False launches all seven every tick
        self.skip_idle_stages = True
        # fused: each tick is one kernel in which every slot runs the seven
        # stages without waiting for the other slots between them (same bits)
        self.fused = bool(fused)
        # with fused and pace_us > 0 a poll is one launch in which every slot
        # runs ticks at its own pace for about pace_us microseconds
        # (ttx_engine_run), instead of `poll` ticks in lockstep -- except in
        # polls whose resident slots all hold the same request at the same
        # step, which the seven tick kernels serve fastest (see _aligned)
        self.pace_us = 0.0
        _check_layout(self.lib)
        self.dev = D.device()
        self.nslots = int(nslots)
        self.rmax = int(rmax)
        if not 1 <= chunk <= 16:
            raise ValueError("chunk (ranks tested per operator application) must be in 1..16")
        self.chunk = int(chunk)
        self.reserve = int(hbm_reserve_bytes)
        if minres_slice < 1:
            raise ValueError("minres_slice must be positive")
        self.minres_slice = int(minres_slice)
        self.host = (Slot * self.nslots)()
        for s in self.host:
            s.status = SLOT_EMPTY
        # numpy views of each descriptor's counter arrays (made once: building
        # an array view of a ctypes array costs more than the sum it feeds)
        self._v_flops = [np.ctypeslib.as_array(s.flops) for s in self.host]
        self._v_clk = [np.ctypeslib.as_array(s.clk) for s in self.host]
        # the descriptors as a (slot, 8-byte field) table, for the stage mask
        self._table = np.frombuffer(self.host, dtype=np.int64).reshape(self.nslots, -1)
        self._col = {nm: getattr(Slot, nm).offset // 8
                     for nm in ("status", "chol", "mixed", "iterative", "sstage", "phase", "k",
                                "sweep")}
        # per slot: the id of the prepared descriptor its request came from
        # (equal ids: equal requests), or a unique negative id
        self._slot_tid = np.zeros(self.nslots, dtype=np.int64)
        self._tids = {}       # descriptor key -> id (ids are never reused)
        self._tid_seq = 0     # last positive id handed out
        self._tid_next = 0    # last negative id handed out
        self.nbytes = ctypes.sizeof(Slot)
        self.dev_slots = torch.zeros(self.nslots * self.nbytes, dtype=torch.uint8, device=self.dev)
        self.pinned = torch.zeros(self.nslots * self.nbytes, dtype=torch.uint8).pin_memory()
        self.slot_req: List[Optional[Request]] = [None] * self.nslots
        self.slot_mem: List[Optional[tuple]] = [None] * self.nslots
        self.arena: List[Optional[torch.Tensor]] = [None] * self.nslots  # reused across requests
        # per-input caches: requests that share A, b, x0 or a seeded z0 reuse
        # the plan, ||b|| and the packed device images of x0 and z0
        self._plans = {}
        self._bnorms = {}
        self._x0_images = {}
        self._z0_images = {}
        # prepared descriptors of admitted requests, keyed by everything the
        # descriptor depends on except the slot's arena (requests that repeat
        # an earlier one are admitted by a copy and four pointers)
        self._templates = {}
        # device work of admissions (arena clears and x0/z0 images), issued in
        # a few multi-tensor launches when the slot array is pushed
        self._dev_zero: List[torch.Tensor] = []
        self._dev_dst: List[torch.Tensor] = []
        self._dev_src: List[torch.Tensor] = []
        self.profiling = False
        self.kernel_ms = np.zeros(NKERNELS)     # per tick kernel (profiling runs only)
        self.kernel_flops = np.zeros(NKERNELS)  # algorithmic flops of retired requests
        self.kernel_launches = 0
        self.paced_polls = 0  # polls served by one paced launch (ttx_engine_run)
        self.minres_iters = 0
        self.lu_solves = 0
        self.chol_fallbacks = 0  # Cholesky solves that found no positive pivot and ran LU
        self.mixed_solves = 0
        self.mixed_refines = 0
        self.mixed_fallbacks = 0
        self.stage_ms = np.zeros(24)  # per-stage device wall clock summed over retired requests
        self._push_all()

    # ------------------------------------------------------------ host <-> device slots
    def _host_bytes(self) -> np.ndarray:
        return np.frombuffer(self.host, dtype=np.uint8)

    def _flush_dev(self) -> None:
        """Issue the admissions' queued arena clears and image copies (on the
        current stream, before the slot descriptors that point at them)."""
        if self._dev_zero:
            torch._foreach_zero_(self._dev_zero)
        if self._dev_dst:
            torch._foreach_copy_(self._dev_dst, self._dev_src)
        self._dev_zero, self._dev_dst, self._dev_src = [], [], []

    def _push_all(self) -> None:
        self._flush_dev()
        self.pinned.numpy()[:] = self._host_bytes()
        self.dev_slots.copy_(self.pinned, non_blocking=False)

    def _push(self, i: int) -> None:
        self._flush_dev()
        a, b = i * self.nbytes, (i + 1) * self.nbytes
        self.pinned.numpy()[a:b] = self._host_bytes()[a:b]
        self.dev_slots[a:b].copy_(self.pinned[a:b], non_blocking=False)

    def _pull(self) -> None:
        self.pinned.copy_(self.dev_slots, non_blocking=False)
        self._host_bytes()[:] = self.pinned.numpy()

    # ------------------------------------------------------------ admission
    def _plan(self, req: Request) -> Slot:
        A, b = req.A, req.b
        d = len(A)
        if d > MAXD:
            raise ValueError(f"at most {MAXD} cores per request")
        if req.precision not in ("fp64", "mixed", "bf16"):
            raise ValueError(f"precision must be 'fp64', 'mixed' or 'bf16', got {req.precision!r}")
        cap = self.capacity(req)
        key = (tuple(A.row_modes), tuple(A.bonds), tuple(b.bonds), req.kickrank,
               req.max_rank, int(req.direct_solve_size), req.precision, cap,
               bool(req.allow_dense_fallback), str(req.local_solver), bool(req.symmetric))
        hit = self._plans.get(key)
        if hit is not None:
            s = Slot()
            ctypes.memmove(ctypes.addressof(s), ctypes.addressof(hit), ctypes.sizeof(Slot))
            return s
        s = Slot()
        s.d = d
        s.kick = req.kickrank
        s.max_rank = 0 if req.max_rank is None else int(req.max_rank)
        s.direct_size = int(req.direct_solve_size)
        s.rmax = cap
        s.chunk = self.chunk
        # mixed and bf16 requests get the scratch of the FP32 operator
        s.mixed = {"fp64": 0, "mixed": 1, "bf16": 2}[req.precision]
        # the explicit-matrix fallback needs the direct solve's buffer for any local size
        s.allow_fallback = 1 if req.allow_dense_fallback else 0
        # the local solvers (common.py:330-360): direct only, MINRES or GMRES above
        # direct_solve_size; GMRES slots get its Krylov basis in their scratch
        ls = str(req.local_solver)
        s.iterative = (0 if ls == "direct" else
                       1 if ls == "minres" or (ls == "auto" and req.symmetric) else 2)
        s.chol = 1 if req.symmetric else 0
        for i in range(d):
            s.n[i] = A.row_modes[i]
        for i in range(d + 1):
            s.ra[i] = A.bonds[i]
            s.rb[i] = b.bonds[i]
        rc = self.lib.ttx_engine_plan(ctypes.byref(s))
        if rc != 0:
            raise ValueError("engine planner rejected the request")
        if len(self._plans) < 256:
            keep = Slot()
            ctypes.memmove(ctypes.addressof(keep), ctypes.addressof(s), ctypes.sizeof(Slot))
            self._plans[key] = keep
        return s

    @staticmethod
    def _train_key(t: TensorTrain):
        """Identity of a train's current contents: object, storage and the
        in-place version counter of every core."""
        return (id(t),) + tuple((c.data_ptr(), c._version) for c in t.cores)

    @staticmethod
    def _cached(cache: dict, key, make):
        hit = cache.get(key)
        if hit is None:
            if len(cache) >= 64:
                cache.pop(next(iter(cache)))
            hit = cache[key] = make()
        return hit[-1]

    def _x0_image(self, x0: TensorTrain, cap: int) -> torch.Tensor:
        """x0's cores at stride ``cap`` in one device buffer (the arena's x block)."""
        def make():
            img = torch.zeros(len(x0) * cap, dtype=D.F64, device=self.dev)
            for k, c in enumerate(x0.cores):
                img[k * cap:k * cap + c.numel()].copy_(c.reshape(-1))
            return (x0, img)
        return self._cached(self._x0_images, (self._train_key(x0), cap), make)

    def _z0_image(self, seed, n: list, kick: int, cap: int) -> torch.Tensor:
        """z0 = rand(modes, kickrank, seed) in the reference's draw order
        (amen.py:80-106), packed at stride ``cap``."""
        d = len(n)

        def make():
            rng = np.random.default_rng(seed)
            zr = [1] + [kick] * (d - 1) + [1]
            img = np.zeros(d * cap)
            for k in range(d):
                core = rng.random((zr[k], n[k], 1, zr[k + 1]))
                img[k * cap:k * cap + core.size] = core.reshape(-1)
            return (D.to_device(img),)
        if seed is None:
            return make()[0]
        return self._cached(self._z0_images, (seed, tuple(n), kick, cap), make)

    def capacity(self, req: Request) -> int:
        """The rank capacity a slot for ``req`` is planned with."""
        return int(req.rank_capacity) if req.rank_capacity is not None else self.rmax

    def _regrow(self, req: Request) -> bool:
        """After ``req`` failed on rank capacity: raise its capacity and reset it
        for another admission; False when it cannot grow."""
        cap = self.capacity(req)
        limit = self.max_rank_capacity
        if req.max_rank is not None:
            limit = min(limit, int(req.max_rank))
        if req.error != _STATUS_TEXT[-1] or cap >= limit:
            return False
        req.rank_capacity = min(2 * cap, limit)
        req.result = req.info = req.error = None
        self.regrown += 1
        return True

    def slot_bytes_needed(self, req: Request) -> int:
        s = self._plan(req)
        d = s.d
        return 8 * (d * s.cap_core + d * s.cap_zcore + s.env_size + s.scr_size)

    def admit(self, i: int, req: Request, push: bool = True) -> None:
        """Place ``req`` in free slot ``i``: arena, descriptor, initial trains.

        With ``push=False`` the descriptor stays on the host until the caller
        pushes the slot array (one copy for a whole admission round); the
        arena's clear and the x0/z0 copies are issued with that push.
        """
        tkey = None
        self._tid_next += 1
        tid = -self._tid_next  # unique unless the request has a prepared descriptor
        if req.x0 is not None and req.seed is not None:
            tkey = (self._train_key(req.A), self._train_key(req.b), self._train_key(req.x0),
                    req.kickrank, req.max_rank, int(req.direct_solve_size), req.precision,
                    self.capacity(req), bool(req.allow_dense_fallback), str(req.local_solver),
                    bool(req.symmetric), req.residual_tol, req.max_sweeps, req.stagnation_factor,
                    req.local_rtol, req.residual_damp, req.error_threshold, req.seed, self.tiles,
                    self.minres_slice)
            tid = self._tids.get(tkey)
            if tid is None:
                if len(self._tids) >= 256:  # forgetting a key only loses alignment
                    self._tids.pop(next(iter(self._tids)))
                self._tid_seq += 1
                tid = self._tids[tkey] = self._tid_seq
            hit = self._templates.get(tkey)
            if hit is not None:
                tmpl, nx, nz, need, x0_img, z0_img = hit
                s = Slot()
                ctypes.memmove(ctypes.addressof(s), ctypes.addressof(tmpl), ctypes.sizeof(Slot))
                self._place(i, req, s, nx, nz, need, x0_img, z0_img, push)
                self._slot_tid[i] = tid
                return
        if req.kickrank <= 0:
            raise ValueError(f"kickrank must be a positive integer, got {req.kickrank}")
        if req.max_rank is not None and req.max_rank <= 0:
            raise ValueError(f"max_rank must be a positive integer, got {req.max_rank}")
        if req.residual_damp < 1.0:
            raise ValueError(f"residual_damp must be >= 1, got {req.residual_damp}")
        A, b = req.A, req.b
        d = len(A)
        st = _SolverSettings.resolve(
            max_sweeps=req.max_sweeps, residual_tol=req.residual_tol,
            stagnation_factor=req.stagnation_factor, local_solver=req.local_solver,
            symmetric=req.symmetric, direct_solve_size=req.direct_solve_size, return_info=True,
            local_rtol=req.local_rtol, allow_dense_fallback=req.allow_dense_fallback,
            truncation_steps=max(2 * (d - 1), 1))
        s = self._plan(req)
        s.tiles = self.tiles
        s.residual_mode = 1 if st.residual_tol is not None else 0
        s.max_sweeps = st.max_sweeps
        s.track = 1 if st.track_residual else 0
        s.residual_tol = st.residual_tol if st.residual_tol is not None else 0.0
        s.trunc_tol = st.truncation_tol if st.truncation_tol is not None else 0.0
        s.local_rtol = st.local_rtol if st.local_rtol is not None else 0.0
        s.damp = req.residual_damp
        s.error_threshold = req.error_threshold
        s.stagnation = st.stagnation_factor if st.stagnation_factor is not None else 0.0
        s.bneg = b.orth_center if b.orth_center is not None else 0
        s.bnorm = (self._cached(self._bnorms, self._train_key(b), lambda: (b, b.tt_norm()))
                   if st.track_residual else 0.0)
        x0 = req.x0
        if x0 is None:
            from .tensor_train import ones
            x0 = ones(b.physical_dims(), ranks=b.bonds)
        if max(x0.bonds) > self.capacity(req):
            raise ValueError(f"x0 bond {max(x0.bonds)} exceeds the slot capacity "
                             f"{self.capacity(req)}")
        # arena: x | z | env | scratch, kept by the slot across requests
        nx, nz = d * s.cap_core, d * s.cap_zcore
        need = nx + nz + s.env_size + s.scr_size
        x0_img = self._x0_image(x0, s.cap_core)
        zn = [int(s.n[k]) for k in range(d)]
        z0_img = self._z0_image(req.seed, zn, req.kickrank, s.cap_zcore)
        for k in range(d):
            s.A[k] = A.cores[k].data_ptr()
            s.b[k] = b.cores[k].data_ptr()
        zr = [1] + [req.kickrank] * (d - 1) + [1]
        for k in range(d + 1):
            s.rx[k] = x0.bonds[k]
            s.rz[k] = zr[k]
        # x0 is right-orthonormalised first unless its centre is already core 0
        if x0.orth_center == 0:
            s.phase, s.k = PH_ZORTH, d - 1
        else:
            s.phase, s.k = PH_XORTH, (x0.orth_center if x0.orth_center is not None else d - 1)
            if s.k < 1:
                s.phase, s.k = PH_ZORTH, d - 1
        s.status = SLOT_RUNNING
        s.sstage = 0
        s.chk_core = -1
        s.minres_slice = self.minres_slice
        if tkey is not None:
            if len(self._templates) >= 64:
                self._templates.pop(next(iter(self._templates)))
            tmpl = Slot()
            ctypes.memmove(ctypes.addressof(tmpl), ctypes.addressof(s), ctypes.sizeof(Slot))
            self._templates[tkey] = (tmpl, nx, nz, need, x0_img, z0_img)
        self._place(i, req, s, nx, nz, need, x0_img, z0_img, push)
        self._slot_tid[i] = tid

    def _place(self, i: int, req: Request, s: Slot, nx: int, nz: int, need: int,
               x0_img: torch.Tensor, z0_img: torch.Tensor, push: bool) -> None:
        """Bind the prepared descriptor ``s`` to slot ``i``'s arena (grown if too
        small) and queue the arena's clear and its x0/z0 images."""
        arena = self.arena[i]
        if arena is None or arena.numel() < need:
            self.arena[i] = None
            arena = self.arena[i] = torch.empty(need, dtype=D.F64, device=self.dev)
        self._dev_zero.append(arena[nx + nz:need])
        self._dev_dst += [arena[:nx], arena[nx:nx + nz]]
        self._dev_src += [x0_img, z0_img]
        base = arena.data_ptr()
        s.x = base
        s.z = base + 8 * nx
        s.env = base + 8 * (nx + nz)
        s.scr = base + 8 * (nx + nz + s.env_size)
        ctypes.memmove(ctypes.addressof(self.host[i]), ctypes.addressof(s), ctypes.sizeof(Slot))
        self.slot_req[i] = req
        self.slot_mem[i] = (req.A, req.b)
        req.admitted = time.perf_counter()
        if push:
            self._push(i)

    # ------------------------------------------------------------ retirement
    def _retire_many(self, slots: Sequence[int]) -> None:
        """Retire the finished ``slots`` (not pushed): the x blocks of the done
        ones leave their arenas in one concatenating copy, each result's cores
        are views of its part of it."""
        blocks = {}
        done = [i for i in slots if self.host[i].status == SLOT_DONE]
        if done:
            sizes = [int(self.host[i].d) * int(self.host[i].cap_core) for i in done]
            big = torch.cat([self.arena[i][:n] for i, n in zip(done, sizes)])
            for i, part in zip(done, torch.split(big, sizes)):
                blocks[i] = part
        for i in slots:
            self._retire(i, push=False, xs=blocks.get(i))

    def _retire(self, i: int, push: bool = True, xs: Optional[torch.Tensor] = None) -> None:
        s = self.host[i]
        req = self.slot_req[i]
        self.kernel_flops += self._v_flops[i]
        self.minres_iters += int(s.minres_iters)
        self.lu_solves += int(s.lu_solves)
        self.chol_fallbacks += int(s.chol_fallbacks)
        self.mixed_solves += int(s.mixed_solves)
        self.mixed_refines += int(s.mixed_refines)
        self.mixed_fallbacks += int(s.mixed_fallbacks)
        self.stage_ms += self._v_clk[i] * 1e-5  # 100 MHz ticks
        if s.status == SLOT_DONE:
            d = s.d
            cap = s.cap_core
            bonds = list(s.rx[:d + 1])
            modes = list(s.n[:d])
            if xs is None:
                xs = self.arena[i][:d * cap].clone()
            # the block holds core k at offset k cap, padded to cap: each core
            # is a contiguous (r_k, n_k, 1, r_k+1) view of it (one as_strided
            # per core: the cheapest view torch builds from Python)
            base, view = xs.storage_offset(), xs.as_strided
            cores = [view((bonds[k], modes[k], 1, bonds[k + 1]),
                          (modes[k] * bonds[k + 1], bonds[k + 1], bonds[k + 1], 1), base + k * cap)
                     for k in range(d)]
            x = TensorTrain._from_device_cores(cores, bonds, modes)
            x.orth_center = 0
            info = SolverInfo()
            nsw = int(s.sweep) + 1
            info.num_sweeps_run = nsw
            if s.track:
                info.residuals = [float(s.residuals[j]) for j in range(min(nsw, MAXSWEEPS))]
            info.max_bonds = [int(s.max_bonds[j]) for j in range(min(nsw, MAXSWEEPS))]
            info.stop_reason = {1: StopReason.CONVERGED, 2: StopReason.STAGNATED}.get(
                int(s.stop_reason), StopReason.MAX_SWEEPS)
            info.local_solves = int(s.local_solves)
            info.local_inexact = int(s.local_inexact)
            info.local_fallbacks = int(s.local_fallbacks)
            req.result, req.info = x, info
        else:
            req.error = _STATUS_TEXT.get(int(s.status), f"status {int(s.status)}")
        req.finished = time.perf_counter()
        s.status = SLOT_EMPTY
        self.slot_req[i] = None
        self.slot_mem[i] = None
        if push:
            self._push(i)

    # ------------------------------------------------------------ loop
    def _stages(self, lockstep: bool = False) -> int:
        """The kernels the next ticks launch, from the host mirror of the slots.

        The FP32 kernel (FP32 Cholesky, FP32 MINRES) runs while a mixed
        request is resident; the LU kernel (LU, GMRES) while a resident
        request solves with it (``symmetric=False``, GMRES) or, as the last
        poll saw, waits on it after a failed Cholesky.  A solve that turns to
        LU between polls waits untouched until the next poll launches the LU
        kernel: the slot's results do not depend on when its kernels run.
        A fused engine's ticks are the one fused kernel, except in the
        lockstep polls of a paced engine (``lockstep``: the seven kernels)."""
        if self.fused and not lockstep:
            return FUSED
        if not self.skip_idle_stages:
            return STAGES_ALL
        t, c = self._table, self._col
        run = t[:, c["status"]] == SLOT_RUNNING
        m = STAGES_ALL & ~(STAGE_CHOL32 | STAGE_LU)
        if np.any(run & (t[:, c["mixed"]] != 0)):
            m |= STAGE_CHOL32
        ss = t[:, c["sstage"]]
        if np.any(run & ((t[:, c["chol"]] == 0) | (t[:, c["iterative"]] == 2)
                         | (ss == _SSTAGE_LU) | (ss == _SSTAGE_GMRES))):
            m |= STAGE_LU
        return m

    def _aligned(self) -> bool:
        """True when every running slot holds the same request (same prepared
        descriptor) at the same step (phase, site, sweep, solve stage), as the
        host mirror last saw them: such slots stay in step, and lockstep ticks
        of the seven kernels serve them faster than paced polls."""
        t, c = self._table, self._col
        run = t[:, c["status"]] == SLOT_RUNNING
        if not np.any(run):
            return True
        ids = self._slot_tid[run]
        if np.any(ids != ids[0]):
            return False
        for nm in ("phase", "k", "sweep", "sstage"):
            v = t[run, c[nm]]
            if np.any(v != v[0]):
                return False
        return True

    def tick(self, n: int = 1) -> None:
        if self.profiling:
            ms = (_f64 * NKERNELS)(*([0.0] * NKERNELS))
            rc = self.lib.ttx_engine_profile(D.stream(), self.dev_slots.data_ptr(), self.nslots, n, ms)
            _lib.check(rc, "ttx_engine_profile")
            self.kernel_ms += np.array(list(ms))
            self.kernel_launches += n
            return
        paced = self.fused and self.pace_us > 0
        if paced and not self._aligned():
            budget = max(1, int(self.pace_us * 100))  # 100 MHz wall clock; never 0 (no limit)
            rc = self.lib.ttx_engine_run(D.stream(), self.dev_slots.data_ptr(), self.nslots, 1 << 40,
                                         budget)
            _lib.check(rc, "ttx_engine_run")
            self.paced_polls += 1
            return
        stages = self._stages(lockstep=paced)
        if self.graphs:
            g = self._graphs.get((n, stages))
            if g is None:
                h = _vp()
                rc = self.lib.ttx_engine_graph_create(self.dev_slots.data_ptr(), self.nslots, n,
                                                      stages, ctypes.byref(h))
                if rc != _lib.TTX_OK:
                    # the same kernels in the same order, launched one by one
                    warnings.warn(f"HIP graph capture of the engine ticks failed (status {rc}); "
                                  "launching the ticks directly", RuntimeWarning)
                    self.graphs = False
                else:
                    g = self._graphs[(n, stages)] = h
            if g is not None:
                _lib.check(self.lib.ttx_engine_graph_launch(D.stream(), g), "ttx_engine_graph_launch")
                return
        rc = self.lib.ttx_engine_tick(D.stream(), self.dev_slots.data_ptr(), self.nslots, n, stages)
        _lib.check(rc, "ttx_engine_tick")

    def close_graphs(self) -> None:
        """Free the recorded tick graphs (also done when the engine is collected)."""
        for g in self._graphs.values():
            self.lib.ttx_engine_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self.close_graphs()
        except Exception:  # interpreter shutdown: the library may be gone already
            pass

    def active(self) -> int:
        return sum(1 for r in self.slot_req if r is not None)

    def run(self, queue: Sequence[Request], poll: int = 8, clock=None, max_ticks: int = 10 ** 7,
            on_done=None) -> List[Request]:
        """Serve ``queue`` (ordered by arrival time) with continuous batching.

        A request is admitted once its arrival time has passed (by ``clock``,
        default: immediately), a slot is free and HBM has room; every ``poll``
        ticks finished requests are retired.
        """
        pending = list(queue)
        done: List[Request] = []
        ticks = 0
        t0 = time.perf_counter()
        now = clock or (lambda: float("inf"))
        while pending or self.active():
            # admit
            free = [i for i in range(self.nslots) if self.slot_req[i] is None]
            admitted = False
            if pending and free and pending[0].arrival <= now():
                free_b, _ = torch.cuda.mem_get_info(self.dev)
                while pending and free and pending[0].arrival <= now():
                    i = free[0]
                    held = self.arena[i].numel() * 8 if self.arena[i] is not None else 0
                    need = max(self.slot_bytes_needed(pending[0]) - held, 0)
                    if free_b - need < self.reserve and self.active():
                        break
                    self.admit(free.pop(0), pending.pop(0), push=False)
                    free_b -= need
                    admitted = True
            if admitted:
                self._push_all()
            if not self.active():
                if pending:
                    time.sleep(max(0.0, min(0.001, pending[0].arrival - now())))
                continue
            self.tick(poll)
            ticks += poll
            self._pull()
            fin = [i for i in range(self.nslots)
                   if self.slot_req[i] is not None and self.host[i].status != SLOT_RUNNING]
            reqs = [self.slot_req[i] for i in fin]
            self._retire_many(fin)
            for req in reqs:
                if self._regrow(req):
                    pending.insert(0, req)
                    continue
                done.append(req)
                if on_done is not None:
                    on_done(req)
            if fin:
                self._push_all()
            if ticks > max_ticks:
                raise RuntimeError("engine did not finish within the tick limit")
        return done


class EnginePool:
    """Several engines served concurrently, each on its own HIP stream.

    Requests of very different sizes go to different engines
    (``route(request) -> index``), whose launches run side by side on the
    GPU: with lockstep ticks a slot doing a long step holds up the rest of its
    engine (paced engines, ``Engine.pace_us``, do not have that problem).  The
    host launches a poll (``poll`` ticks, or one paced launch) plus an
    asynchronous copy of the slot descriptors on an engine's stream and
    services whichever engine's copy has landed first: retire the finished
    requests, admit arrived ones into free slots (within free HBM), launch the
    next poll.  Every operation of an engine runs on its stream.
    """

    def __init__(self, engines: Sequence[Engine], route):
        self.engines = list(engines)
        self.route = route
        self.streams = [torch.cuda.Stream(device=e.dev) for e in self.engines]
        self.events = [None] * len(self.engines)

    def _launch(self, i: int, poll: int) -> None:
        e, st = self.engines[i], self.streams[i]
        with torch.cuda.stream(st):
            e.tick(poll)
            e.pinned.copy_(e.dev_slots, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        self.events[i] = ev

    def run(self, queue: Sequence[Request], poll: int = 8, clock=None, on_done=None,
            max_cycles: int = 10 ** 7) -> List[Request]:
        """Serve ``queue`` (ordered by arrival); returns the requests as they finish."""
        pending = [list() for _ in self.engines]
        for r in queue:
            pending[self.route(r)].append(r)
        done: List[Request] = []
        now = clock or (lambda: float("inf"))
        cycles = 0
        while any(pending) or any(ev is not None for ev in self.events):
            progressed = False
            for i, e in enumerate(self.engines):
                ev = self.events[i]
                if ev is not None:
                    if not ev.query():
                        continue
                    self.events[i] = None
                    with torch.cuda.stream(self.streams[i]):
                        e._host_bytes()[:] = e.pinned.numpy()
                        fin = [j for j in range(e.nslots)
                               if e.slot_req[j] is not None and e.host[j].status != SLOT_RUNNING]
                        reqs = [e.slot_req[j] for j in fin]
                        e._retire_many(fin)
                        for req in reqs:
                            if e._regrow(req):
                                pending[i].insert(0, req)
                                continue
                            done.append(req)
                            if on_done is not None:
                                on_done(req)
                        if fin:
                            e._push_all()
                    progressed = True
                with torch.cuda.stream(self.streams[i]):
                    q = pending[i]
                    if q and q[0].arrival <= now():
                        free = [j for j in range(e.nslots) if e.slot_req[j] is None]
                        admitted = False
                        if free:
                            free_b, _ = torch.cuda.mem_get_info(e.dev)
                            while q and free and q[0].arrival <= now():
                                j = free[0]
                                held = e.arena[j].numel() * 8 if e.arena[j] is not None else 0
                                need = max(e.slot_bytes_needed(q[0]) - held, 0)
                                if free_b - need < e.reserve and e.active():
                                    break
                                e.admit(free.pop(0), q.pop(0), push=False)
                                free_b -= need
                                admitted = True
                        if admitted:
                            e._push_all()
                    if e.active():
                        self._launch(i, poll)
                        progressed = True
            cycles += 1
            if cycles > max_cycles:
                raise RuntimeError("engine pool did not finish within the cycle limit")
            if not progressed:
                time.sleep(0.0002)
        return done


def amen_batched(requests: Sequence[Request], nslots: Optional[int] = None, rmax: int = 64,
                 poll: int = 8) -> List[Request]:
    """Solve all ``requests`` on the engine; returns them with ``result``/``info`` set."""
    eng = Engine(nslots or len(requests), rmax)
    return eng.run(list(requests), poll=poll)
