"""Host-side launch decisions of the engine (CPU): which kernels a tick
launches (stage masks) and when a paced engine serves a poll with lockstep
ticks.  Both are functions of the host mirror of the slot table only, so they
run here on stub engines whose tables are written by hand."""

import ctypes
import types

import numpy as np

from ttnumpy_amd.engine import (FUSED, SLOT_DONE, SLOT_EMPTY, SLOT_RUNNING, STAGE_CHOL32, STAGE_LU,
                                STAGES_ALL, Engine, Slot)

COLS = ("status", "chol", "mixed", "iterative", "sstage", "phase", "k", "sweep")
BASE = STAGES_ALL & ~(STAGE_CHOL32 | STAGE_LU)


def _stub(n, fused=False, skip=True):
    """An object with the fields Engine._stages and Engine._aligned read: n
    running, symmetric, FP64, direct-solve slots of one prepared descriptor,
    all at the same step."""
    host = (Slot * n)()
    e = types.SimpleNamespace()
    e.host = host
    e._table = np.frombuffer(host, dtype=np.int64).reshape(n, -1)
    e._col = {nm: getattr(Slot, nm).offset // 8 for nm in COLS}
    e._slot_tid = np.ones(n, dtype=np.int64)
    e.fused, e.skip_idle_stages = fused, skip
    for s in host:
        s.status, s.chol, s.mixed, s.iterative, s.sstage = SLOT_RUNNING, 1, 0, 1, 0
        s.phase, s.k, s.sweep = 5, 3, 1
    return e


def stages(e, lockstep=False):
    return Engine._stages(e, lockstep)


def aligned(e):
    return Engine._aligned(e)


def test_table_view_reads_the_slot_fields():
    e = _stub(3)
    e.host[1].sweep = 7
    e.host[2].phase = 9
    assert e._table[1, e._col["sweep"]] == 7 and e._table[2, e._col["phase"]] == 9
    assert ctypes.sizeof(Slot) % 8 == 0


def test_stage_masks():
    e = _stub(4)
    assert stages(e) == BASE                          # FP64 symmetric: no FP32, no LU kernel
    e.host[2].mixed = 1
    assert stages(e) == BASE | STAGE_CHOL32           # a mixed request is resident
    e.host[2].status = SLOT_DONE
    assert stages(e) == BASE                          # ... but no longer running
    e.host[0].chol = 0
    assert stages(e) == BASE | STAGE_LU               # symmetric=False solves with LU
    e.host[0].chol = 1
    e.host[3].iterative = 2
    assert stages(e) == BASE | STAGE_LU               # GMRES runs in the LU kernel
    e.host[3].iterative = 1
    e.host[1].sstage = 2
    assert stages(e) == BASE | STAGE_LU               # a failed Cholesky waits on LU
    e.host[1].sstage = 6
    assert stages(e) == BASE | STAGE_LU               # a GMRES solve in flight
    e.host[1].status = SLOT_EMPTY
    assert stages(e) == BASE                          # an empty slot's stale fields count for nothing
    e.skip_idle_stages = False
    assert stages(e) == STAGES_ALL
    e.fused = True
    assert stages(e) == FUSED
    assert stages(e, lockstep=True) == STAGES_ALL     # a paced engine's lockstep polls


def test_aligned_polls():
    e = _stub(5)
    assert aligned(e)
    e.host[4].sweep = 2                               # one slot a sweep ahead
    assert not aligned(e)
    e.host[4].status = SLOT_DONE                      # ... finished: the rest are in step
    assert aligned(e)
    e.host[3].sstage = 3                              # same site, MINRES still running
    assert not aligned(e)
    e.host[3].sstage = 0
    e._slot_tid[1] = 2                                # another request at the same step
    assert not aligned(e)
    e._slot_tid[:] = -np.arange(1, 6)                 # requests without a prepared descriptor
    assert not aligned(e)
    for s in e.host[1:]:
        s.status = SLOT_EMPTY
    assert aligned(e)                                 # one running slot is always in step
    e.host[0].status = SLOT_EMPTY
    assert aligned(e)                                 # nothing running
