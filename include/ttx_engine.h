// Slot descriptor of the batched AMEn engine (shared by the kernels and the
// host planner).  Every field is 8 bytes wide so the host mirror of the
// layout is the plain sequence of fields.
#pragma once

#include <stdint.h>

#include "ttx.h"

#define TTX_MAXD 64
#define TTX_MAXSWEEPS 64
#define TTX_ENGINE_KERNELS 7  /* per tick: env, solve, chol32, chol, lu, post, check */
/* stage masks of a tick: bit i launches kernel i of the list above */
#define TTX_ENGINE_STAGES_ALL 0x7f
#define TTX_ENGINE_STAGE_CHOL32 0x04 /* FP32 Cholesky and FP32 MINRES (mixed requests) */
#define TTX_ENGINE_STAGE_LU 0x10     /* LU and GMRES */
/* one kernel per tick that runs all seven stages for each slot, the slots
 * not waiting for each other between stages (same bits as the seven) */
#define TTX_ENGINE_FUSED 0x80

enum ttx_phase {
  TTX_PH_XORTH = 1,     // right-orthonormalise x0, k = d-1 .. 1
  TTX_PH_ZORTH = 2,     // right-orthonormalise the random z0, k = d-1 .. 1
  TTX_PH_FIT_INIT = 3,  // right environments of the z fit, k = d-1 .. 0
  TTX_PH_FIT_FWD = 4,   // z fit, forward, k = 0 .. d-2
  TTX_PH_FIT_LAST = 5,  // left z environments at d-1
  TTX_PH_FIT_BWD = 6,   // z fit, backward, k = d-1 .. 0
  TTX_PH_INIT_R = 7,    // all right environments, k = d-1 .. 0
  TTX_PH_FWD = 8,       // AMEn forward half-sweep, k = 0 .. d-2
  TTX_PH_LAST = 9,      // left environments at d-1
  TTX_PH_BWD = 10,      // AMEn backward half-sweep, k = d-1 .. 0
  TTX_PH_CHECK = 11,    // per-sweep residual and stopping test
  TTX_PH_RED_L = 12,    // redundant-bond reduction, left QR sweep
  TTX_PH_RED_R = 13,    // redundant-bond reduction, right QR sweep
  TTX_PH_DONE = 14
};

enum ttx_slot_status {
  TTX_SLOT_RUNNING = 0,
  TTX_SLOT_DONE = 1,
  TTX_SLOT_ERR_CAPACITY = -1,  // a bond outgrew the slot's rank capacity
  TTX_SLOT_ERR_NONFINITE = -2,
  TTX_SLOT_ERR_SINGULAR = -3,
  TTX_SLOT_EMPTY = -100
};

// scratch regions (indices into scr_off)
enum {
  TTX_SCR_U = 0, TTX_SCR_BLOC, TTX_SCR_TMP, TTX_SCR_ZNEW, TTX_SCR_MAT, TTX_SCR_LU, TTX_SCR_MR,
  TTX_SCR_S1, TTX_SCR_S2, TTX_SCR_SVDW, TTX_SCR_USVD, TTX_SCR_SSVD, TTX_SCR_VSVD, TTX_SCR_ENR,
  TTX_SCR_CHK, TTX_SCR_ENVS, TTX_SCR_SCAL, TTX_SCR_M32, TTX_SCR_GMRES, TTX_SCR_COUNT
};

typedef struct ttx_slot {
  // ---- problem (host-written)
  int64_t d, kick, max_rank, direct_size, iterative, residual_mode, max_sweeps, track;
  int64_t rmax, nmax, ramax, rbmax, bneg, allow_fallback, chunk;
  // (iterative: 0 direct local solves only, 1 MINRES, 2 GMRES above direct_size)
  int64_t chol;  // 1: direct solves try Cholesky first (symmetric problems), LU if not SPD
  int64_t mixed; // 1: direct Cholesky solves of size >= 256 factor in FP32 with FP64 refinement,
                 //    MINRES solves run on the FP32 operator with FP64 defect correction;
                 // 2: the same with the MINRES operator's products on BF16 operands
  int64_t minres_slice; // MINRES iterations per tick (a long local solve spans ticks)
  int64_t n[TTX_MAXD];
  int64_t ra[TTX_MAXD + 1];
  int64_t rb[TTX_MAXD + 1];
  double residual_tol, trunc_tol, local_rtol, damp, error_threshold, stagnation, bnorm, pad1;
  // ---- layout (planner-written, doubles)
  int64_t cap_core, cap_zcore, cap_loc, direct_cap, lmax, scr_size, env_size;
  int64_t tiles;  // tile variants (host-written): bit 0, Cholesky rows below the diagonal in 64-row tiles
  int64_t cap_env[5], pad3[3];
  int64_t env_off[10], pad4[6];
  int64_t scr_off[TTX_SCR_COUNT], pad5[32 - TTX_SCR_COUNT];
  const double* A[TTX_MAXD];
  const double* b[TTX_MAXD];
  double* x;
  double* z;
  double* env;
  double* scr;
  // ---- state (device-written, host-initialised)
  int64_t rx[TTX_MAXD + 1];
  int64_t rz[TTX_MAXD + 1];
  int64_t phase, k, sweep, status, stop_reason, local_solves, local_inexact, local_fallbacks;
  int64_t rank_decisions, rank_capped, red_last, svd_noconv, ticks, minres_iters, lu_solves,
      chol_fallbacks;
  int64_t mixed_solves, mixed_refines, mixed_fallbacks;
  int64_t sstage;   // local solve in flight: 1 Cholesky, 2 LU, 4 FP32 Cholesky (this tick), 3 MINRES (resumes)
  int64_t chk_core; // residual check in flight: next core (-1: not started)
  int64_t chk_cdim; // and the width of its carried R factor
  double flops[TTX_ENGINE_KERNELS];  // algorithmic flops banked per kernel
  double residuals[TTX_MAXSWEEPS];
  int64_t max_bonds[TTX_MAXSWEEPS];
  // wall-clock ticks (100 MHz) spent per stage, thread 0's view:
  // 0 local rhs, 1 explicit matrix, 2 LU, 3 MINRES, 4 exact residual, 5 post, 6 check, 7 env,
  // inside post (sweep steps only): 8 residual core, 9 SVD, 10 rank choice, 11 enrichment
  // and QR, 12 z shift; inside the LU: 13 panels, 14 in-panel updates, 15 block rows,
  // 16 trailing updates, 17 back substitution (inside the Cholesky: 13 diagonal-block
  // factorizations, 14 diagonal-block updates, 16 rows below them, 17 back
  // substitution); mixed direct solves: 18 FP32 factor, 19 refinement; inside FP64
  // MINRES: 20 operator applications, 21 vector updates and reductions; 22-23 spare
  int64_t clk[24];
} ttx_slot;

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Batched device-resident AMEn (the sweep of ttnumpy/solvers/amen.py:395-725
 * with the residual check of ttnumpy/solvers/common.py:118-146), one slot per
 * request.  The reference runs one solve per Python call; these entry points
 * serve many at once and have no single counterpart there.
 */

/* sizeof(ttx_slot), for the host mirror's layout check. */
int64_t ttx_engine_slot_bytes(void);
/* byte offset of a field: 0 n, 1 residual_tol, 2 cap_core, 3 A, 4 rx, 5 residuals, 6 max_bonds. */
int64_t ttx_engine_field_offset(int which);
/*
 * Fill the layout block of a host-side slot whose problem block is set:
 * per-core, environment and scratch capacities and offsets (in doubles) of the
 * slot's arena for rank capacity rmax.  Returns TTX_EINVAL on bad sizes.
 */
int ttx_engine_plan(ttx_slot* s);
/*
 * Advance every slot of the device array `slots` by `ticks` ticks (seven
 * kernels each: environments; local rhs and MINRES; FP32 Cholesky with FP64
 * refinement; Cholesky; LU; truncation/enrichment/shift; residual check).  Slots with status != RUNNING are skipped.  ticks == 0
 * only prepares the kernels' LDS attributes.  `stages` (a subset of
 * TTX_ENGINE_STAGES_ALL, else TTX_EINVAL; or TTX_ENGINE_FUSED, which
 * launches the one fused kernel instead) names the kernels a tick launches:
 * a slot whose local solve waits on a kernel left out stays where it is,
 * untouched, until a tick launches that kernel, so leaving out kernels no
 * slot needs changes no result.
 */
int ttx_engine_tick(ttx_stream_t stream, ttx_slot* slots, int64_t nslots, int64_t ticks,
                    uint32_t stages);
/*
 * Ticks of every slot in one launch, each slot at its own pace: the
 * workgroup of a slot runs the seven stages of a tick in order, pass after
 * pass, and stops when its slot is no longer running, after `passes` passes,
 * or (budget > 0) at the first pass boundary `budget` ticks of the 100 MHz
 * wall clock after it started (one pass at least).  A slot's results are
 * those of as many ttx_engine_tick ticks, bit for bit.
 */
int ttx_engine_run(ttx_stream_t stream, ttx_slot* slots, int64_t nslots, int64_t passes,
                   int64_t budget);
/*
 * ttx_engine_tick (all seven kernels) with HIP events around each kernel and a
 * synchronisation per tick; ms_out[0..6] accumulate the kernels' milliseconds.
 */
int ttx_engine_profile(ttx_stream_t stream, ttx_slot* slots, int64_t nslots, int64_t ticks,
                       double* ms_out);
/*
 * `ticks` ticks of ttx_engine_tick with the kernels of `stages` recorded once
 * as a HIP graph (captured on
 * a private stream; *graph_out receives an opaque handle).  The kernels read
 * everything through the device slot table `slots`, which must outlive the
 * graph; admissions and retirements change the table's contents, not its
 * address, so the graph serves every later poll.  launch replays it on
 * `stream`; destroy frees it.
 */
int ttx_engine_graph_create(ttx_slot* slots, int64_t nslots, int64_t ticks, uint32_t stages,
                            void** graph_out);
int ttx_engine_graph_launch(ttx_stream_t stream, void* graph);
int ttx_engine_graph_destroy(void* graph);

#ifdef __cplusplus
}
#endif
