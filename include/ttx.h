/*
 * ttx -- C ABI of the MI355X (gfx950) tensor-train solver kernels.
 *
 * Every entry point takes device pointers (FP64, element strides in units of
 * elements, int64 sizes), a HIP stream, and returns a status code.  No entry
 * point allocates, frees or retains a caller buffer; scratch memory comes in
 * through an explicit `work` argument whose size the matching *_work_size
 * query returns (in doubles).
 *
 * The reference library (ttnumpy 0.3.0) is pure Python over NumPy/SciPy and
 * has no FFI of its own; each entry below names the reference operation it
 * replaces (file:line in the reference tree).  The Python package
 * `ttnumpy_amd` binds these through ctypes and exposes the reference's
 * Python API (`amen`, `als`, `TensorTrain.truncate`, `SolverInfo`).
 */
#ifndef TTX_H_
#define TTX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* ttx_stream_t; /* hipStream_t */

enum {
  TTX_OK = 0,
  TTX_EINVAL = 1,      /* argument error            -> ValueError            */
  TTX_ENONFINITE = 2,  /* NaN/inf input             -> ValueError("Core i has NaN or infinite entries") */
  TTX_ENOCONV = 3,     /* iteration did not converge -> warning + retry      */
  TTX_EHIP = 4,        /* HIP launch/runtime error   -> RuntimeError          */
  TTX_ESINGULAR = 5,   /* exactly singular pivot     -> numpy LinAlgError    */
  TTX_ENOTSPD = 6,     /* Cholesky met a pivot <= 0 (or NaN): not positive definite */
  TTX_ESCRATCH = 7     /* the call needs more than a fixed scratch array holds -> ValueError */
};

/* Library version (major*10000 + minor*100 + patch). */
int ttx_version(void);

/*
 * Strided batched FP64 GEMM on MFMA (v_mfma_f64_16x16x4f64):
 *   C[b](m,n) = alpha * sum_k A[b](m,k) * B[b](k,n) + beta * C[b](m,n)
 * with A[b](m,k) = A[b*a_bs + m*a_ms + k*a_ks] etc.  beta == 0 never reads C.
 * Replaces every np.tensordot / @ / einsum of the hot path:
 *   environment updates  ttnumpy/solvers/common.py:178-217
 *   one-site local apply  ttnumpy/solvers/one_site.py:98-108
 *   residual core         ttnumpy/solvers/amen.py:142-159
 *   core absorb           ttnumpy/utils.py:72-83
 *   TT matvec cores       ttnumpy/tensor_train.py:483-495
 */
int ttx_gemm_f64(ttx_stream_t stream, int64_t batch, int64_t M, int64_t N, int64_t K,
                 double alpha,
                 const double* A, int64_t a_ms, int64_t a_ks, int64_t a_bs,
                 const double* B, int64_t b_ks, int64_t b_ns, int64_t b_bs,
                 double beta,
                 double* C, int64_t c_ms, int64_t c_ns, int64_t c_bs);

/*
 * ttx_gemm_f64's contraction in FP32: A / a_scale and B / b_scale are rounded
 * to FP32, products accumulate in FP32 on v_mfma_f32_16x16x4f32, and
 * C = alpha * a_scale * b_scale * (product) + beta * C in FP64.  The scales
 * are device scalars (e.g. the operands' norms, so that a train's norm does
 * not leave FP32's range); NULL, zero or non-finite means 1.  The "FP32
 * contraction" of the mixed rounding form of TensorTrain.truncate (the
 * absorb products of ttnumpy/utils.py:72-83).
 */
int ttx_gemm_f32acc(ttx_stream_t stream, int64_t batch, int64_t M, int64_t N, int64_t K,
                    double alpha, const double* A, int64_t a_ms, int64_t a_ks, int64_t a_bs,
                    const double* B, int64_t b_ks, int64_t b_ns, int64_t b_bs, double beta,
                    double* C, int64_t c_ms, int64_t c_ns, int64_t c_bs, const double* a_scale,
                    const double* b_scale);

/*
 * dst[i0..i5] = alpha * src[i0..i5] over an ndim-dimensional box (ndim <= 6),
 * with independent element strides.  Permutes, slices, block placement and
 * zero padding of cores (ttnumpy/tensor_train.py:254-322,
 * ttnumpy/solvers/amen.py:348-356, one_site.py:75 transpose).
 */
int ttx_copy_f64(ttx_stream_t stream, int ndim, const int64_t* dims,
                 const double* src, const int64_t* src_strides,
                 double* dst, const int64_t* dst_strides, double alpha);

/* dst[0..n) = value. */
int ttx_fill_f64(ttx_stream_t stream, double* dst, int64_t n, double value);

/* z = a*x + b*y (z may alias x or y). */
int ttx_axpby_f64(ttx_stream_t stream, int64_t n, double a, const double* x, double b,
                  const double* y, double* z);

/* out[0] = ||x||_2 (deterministic).  np.linalg.norm at common.py:787,852, amen.py:250. */
int ttx_nrm2_f64(ttx_stream_t stream, int64_t n, const double* x, double* out);

/* out[0] = ||a - b||_2 (deterministic). */
int ttx_diff_nrm2_f64(ttx_stream_t stream, int64_t n, const double* a, const double* b, double* out);

/*
 * Economic Householder QR of the m x n matrix A(i,j) = A[i*a_rs + j*a_cs]:
 * Q (m x k) and R (k x n), k = min(m, n), written with their own strides.
 * scipy.linalg.qr(mode="economic") at ttnumpy/utils.py:101,130 and
 * ttnumpy/solvers/one_site.py:188,204.  Q may be NULL (only R is formed, as the
 * norm sweep of ttnumpy/tensor_train.py:509-540 needs).  Large matrices
 * (min(m, n) >= 64, m n >= 2^17) take LAPACK's blocked scheme with the
 * trailing updates on the GEMM across the chip.  Work: ttx_qr_work_size(m, n).
 */
int64_t ttx_qr_work_size(int64_t m, int64_t n);
int ttx_qr_f64(ttx_stream_t stream, int64_t m, int64_t n,
               const double* A, int64_t a_rs, int64_t a_cs,
               double* Q, int64_t q_rs, int64_t q_cs,
               double* R, int64_t r_rs, int64_t r_cs, double* work);

/*
 * Thin SVD A = U diag(S) Vt of the m x n matrix A (strided), k = min(m, n),
 * singular values descending.  One-sided Jacobi (after a QR of the tall
 * side).  scipy.linalg.svd at ttnumpy/utils.py:347,359 (gesdd/gesvd) and
 * ttnumpy/solvers/amen.py:254,301 (gesvd).  info[0] receives TTX_OK,
 * TTX_ENOCONV (sweep limit) or TTX_ENONFINITE.  Work: ttx_svd_work_size.
 * With k >= 139 the Jacobi runs across the chip, one launch per round, and
 * the call synchronizes the stream once per sweep (and once after checking
 * the input for NaN/inf): such calls cannot be captured in a HIP graph.
 */
int64_t ttx_svd_work_size(int64_t m, int64_t n);
int ttx_svd_f64(ttx_stream_t stream, int64_t m, int64_t n,
                const double* A, int64_t a_rs, int64_t a_cs,
                double* U, int64_t u_rs, int64_t u_cs, double* S,
                double* Vt, int64_t v_rs, int64_t v_cs, double* work, int32_t* info);

/*
 * ttx_svd_f64 without a work argument and without any device allocation.
 * No counterpart in the reference (whose SVD is LAPACK's, which allocates
 * its own work): this entry is for callers that cannot allocate device
 * memory, such as code under HIP graph capture or a serving loop that
 * forbids hipMalloc.  The work buffer is a fixed, file-scope device array of
 * ttx_svd_noalloc_max_work() doubles (32 MiB; each device holds its own copy
 * until the process ends).  Shapes whose ttx_svd_work_size exceeds it return
 * TTX_ESCRATCH and touch nothing.  Results are those of ttx_svd_f64 on the
 * same input.  All calls on one device share the array, so calls that
 * overlap on one device (two streams, two host threads) race: serialize them.
 */
int64_t ttx_svd_noalloc_max_work(void);
int ttx_svd_f64_noalloc(ttx_stream_t stream, int64_t m, int64_t n,
                        const double* A, int64_t a_rs, int64_t a_cs,
                        double* U, int64_t u_rs, int64_t u_cs, double* S,
                        double* Vt, int64_t v_rs, int64_t v_cs, int32_t* info);

/*
 * Solve M u = b for a dense n x n system by LU with partial pivoting.
 * `Mb` holds [M | b] row-major (n rows, n+1 columns, row stride ld >= n+1)
 * and is destroyed; `u` receives the solution.  scipy.linalg.solve at
 * ttnumpy/solvers/common.py:767,802 (getrf/getrs).  info[0]: TTX_OK or
 * TTX_ESINGULAR.  One workgroup with the row permutation and the panels in
 * its 150 KiB of LDS: n <= 12800 (TTX_EINVAL above).  Work:
 * ttx_lu_work_size(n) doubles (reserved, may be NULL).
 */
int64_t ttx_lu_work_size(int64_t n);
/*
 * Solve M u = b for a symmetric positive definite n x n system by Cholesky
 * (M = L L^T from the lower triangle; the upper triangle is never read).
 * Layout as ttx_lu_solve_f64: [M | b] row-major with row stride ld >= n+1,
 * destroyed.  info[0]: TTX_OK, or TTX_ENOTSPD when a pivot is <= 0 or not
 * finite (u is then undefined and Mb partly overwritten: the caller rebuilds
 * it and solves with ttx_lu_solve_f64).  The direct local solve of
 * ttnumpy/solvers/common.py:767 for symmetric problems; scipy.linalg.solve
 * computes the same solution by getrf/getrs (or sysv when it detects exact
 * symmetry), equal to this one up to rounding.
 */
int ttx_chol_solve_f64(ttx_stream_t stream, int64_t n, double* Mb, int64_t ld, double* u,
                       int32_t* info);
int ttx_lu_solve_f64(ttx_stream_t stream, int64_t n, double* Mb, int64_t ld, double* u,
                     double* work, int32_t* info);

/*
 * MINRES (the recurrences and stopping tests of scipy.sparse.linalg.minres,
 * called at ttnumpy/solvers/common.py:784) with the operator applied by the
 * caller between steps.  State lives in `st` (TTX_MINRES_STATE doubles) and
 * the vectors in `vec` (8 * n doubles: x, v, y, r1, r2, w, w1, w2).
 *   init : x = x0 (or 0), r1 = b - A x0 given `ax0` = A x0 (or NULL), sets v
 *   step : given y_in = A v, one iteration; writes the next v
 * st[0] = istop, st[1] = itn; the host polls them.  See krylov.hip.
 */
#define TTX_MINRES_STATE 32
int ttx_minres_init_f64(ttx_stream_t stream, int64_t n, const double* b, const double* x0,
                        const double* ax0, double rtol, int64_t maxiter, double* vec, double* st);
int ttx_minres_step_f64(ttx_stream_t stream, int64_t n, const double* y_in, double* vec,
                        double* st);

/*
 * Restarted GMRES (the iteration of scipy.sparse.linalg.gmres with atol = 0,
 * no preconditioner, restart m = min(20, n), maxiter = 10 n cycles, called at
 * ttnumpy/solvers/common.py:786): modified Gram-Schmidt Arnoldi, LAPACK-style
 * Givens rotations, the adaptive inner tolerance, the true residual after each
 * cycle.  `work` holds ttx_gmres_work_size(n, m) doubles (the Krylov basis V,
 * x at V + (m+1) n, A x at V + (m+2) n, ...), `st` TTX_GMRES_STATE doubles:
 * st[0] done, st[2] info (0 or maxiter), st[10] |b - A x|, st[13] inner steps.
 *   init    : x = x0 (or 0); `ax0` = A x0 when x0 is given, else NULL
 *   arnoldi : step `col` of a cycle, after the caller wrote A V[col] into V[col+1]
 *   finish  : end of a cycle, x += y V
 *   check   : after the caller wrote A x into V + (m+2) n; next cycle or done
 * Steps after the inner stop test (st[1] set) or after done leave the state alone.
 */
#define TTX_GMRES_STATE 16
int64_t ttx_gmres_work_size(int64_t n, int64_t m);
int ttx_gmres_init_f64(ttx_stream_t stream, int64_t n, int64_t m, const double* b,
                       const double* x0, const double* ax0, double rtol, int64_t maxiter,
                       double* work, double* st);
int ttx_gmres_arnoldi_f64(ttx_stream_t stream, int64_t n, int64_t m, int64_t col, double* work,
                          double* st);
int ttx_gmres_finish_f64(ttx_stream_t stream, int64_t n, int64_t m, double* work, double* st);
int ttx_gmres_check_f64(ttx_stream_t stream, int64_t n, int64_t m, const double* b, double* work,
                        double* st);

/*
 * The one-site local matvec (ttnumpy/solvers/one_site.py:78-115, the
 * LinearOperator of the local system): for B column vectors
 *   out[z, fl, i, fr] = alpha * sum psi[tl, a, fl] A[a, i, j, b] y[z, tl, j, tr] phi[tr, b, fr]
 *                       + beta * out[z, fl, i, fr]
 * with psi (rl, Ra, fl), A (Ra, nr, nc, Rb), phi (rr, Rb, fr), y (B, rl, nc, rr)
 * and out (B, fl, nr, fr), as three strided MFMA GEMMs through the scratch
 * s1 (B rl nc Rb fr doubles) and s2 (B rl Ra nr fr doubles).
 */
int ttx_local_apply_f64(ttx_stream_t stream, const double* psi, int64_t rl, int64_t Ra,
                        int64_t fl, const double* A, int64_t nr, int64_t nc, int64_t Rb,
                        const double* phi, int64_t rr, int64_t fr, const double* y, int64_t B,
                        double alpha, double beta, double* out, double* s1, double* s2);

/*
 * `count` MINRES iterations of the local system (the minres call at
 * ttnumpy/solvers/common.py:784), issued from C: each applies the local
 * matvec (as ttx_local_apply_f64, B = 1) to v = vec + n into y_in, then runs
 * ttx_minres_step_f64.  n = fl nr fr must equal rl nc rr.  Steps after the
 * stop test has fired leave the state unchanged.
 */
int ttx_minres_local_steps_f64(ttx_stream_t stream, const double* psi, int64_t rl, int64_t Ra,
                               int64_t fl, const double* A, int64_t nr, int64_t nc, int64_t Rb,
                               const double* phi, int64_t rr, int64_t fr, double* s1, double* s2,
                               double* y_in, double* vec, double* st, int64_t count);

/*
 * One GMRES cycle of the local system issued from C: m arnoldi steps, each
 * after the local matvec of V[col] (skipped once the inner loop has stopped),
 * then finish, the matvec of x and check.
 */
int ttx_gmres_local_cycle_f64(ttx_stream_t stream, const double* psi, int64_t rl, int64_t Ra,
                              int64_t fl, const double* A, int64_t nr, int64_t nc, int64_t Rb,
                              const double* phi, int64_t rr, int64_t fr, double* s1, double* s2,
                              int64_t m, const double* b, double* work, double* st);

/*
 * Rank rules.
 * truncated_bond: ttnumpy/utils.py:22-69 -> out[0] = bond.
 */
int ttx_truncated_bond_f64(ttx_stream_t stream, const double* S, int64_t len, double delta,
                           int64_t max_rank, int32_t* out);
/*
 * Residual-budget rank (ttnumpy/solvers/common.py:812-866): given
 * contrib (N x limit, column-major, column r = A_loc vec(s_r u_r v_r^T)),
 * b_loc, the singular values S (S[0] == 0 gives rank 1), the scalars
 * exact_residual = *exact (device) and half_trunc_tol = 0.5*truncation_tol,
 * damp: out[0] = rank, out[1] = 1 if the budget was never met.  Work: N doubles.
 */
int ttx_rank_by_residual_f64(ttx_stream_t stream, int64_t N, int64_t limit, const double* S,
                             const double* contrib, const double* b_loc, const double* exact,
                             double damp, double half_trunc_tol, double* work, int32_t* out);
/*
 * terms (N x limit, column-major, N = p*q): column r = vec(S[r] * U[:, r] * Vt[r, :])
 * with the row index (i, j) -> i*q + j, i < p, j < q (common.py:855-857).
 */
int ttx_svd_terms_f64(ttx_stream_t stream, int64_t p, int64_t q, int64_t limit,
                      const double* U, int64_t u_rs, int64_t u_cs, const double* S,
                      const double* Vt, int64_t v_rs, int64_t v_cs, double* terms);

/*
 * Explicit one-site local matrix into out[row*o_rs + col*o_cs]:
 *   M[(f_l,i,f_r), (t_l,j,t_r)] = sum_{a,b} psi[t_l,a,f_l] A[a,i,j,b] phi[t_r,b,f_r]
 * psi (rl, Ra, fl), A (Ra, nr, nc, Rb), phi (rr, Rb, fr), all contiguous.
 * ttnumpy/solvers/one_site.py:49-75 (einsum + transpose).
 */
int ttx_local_matrix_f64(ttx_stream_t stream, const double* psi, int64_t rl, int64_t Ra,
                         int64_t fl, const double* A, int64_t nr, int64_t nc, int64_t Rb,
                         const double* phi, int64_t rr, int64_t fr, double* out, int64_t o_rs,
                         int64_t o_cs);

/*
 * One core of the TTM @ TT product (ttnumpy/tensor_train.py:483-495):
 *   out[(a*rx + m)*o_l + i*o_i + (b*rx2 + n)*o_r] = sum_j A[a,i,j,b] x[m,j,n]
 */
int ttx_ttm_core_f64(ttx_stream_t stream, const double* A, int64_t Ra, int64_t nr, int64_t nc,
                     int64_t Rb, const double* x, int64_t rx, int64_t rx2, double* out,
                     int64_t o_l, int64_t o_i, int64_t o_r);

/* X[r*rs + c*cs] *= s[c] (U*S absorbs, ttnumpy/utils.py:400-404,449-453). */
int ttx_scale_cols_f64(ttx_stream_t stream, int64_t rows, int64_t cols, double* X, int64_t rs,
                       int64_t cs, const double* s);

/* out[(a m), i, j, (b n)] = A[a, i, j, b] * B[m, i, j, n] for cores A (ra, ni, nj, rb)
 * and B (sa, ni, nj, sb): one core of the element-wise product of two trains
 * (ttnumpy/tensor_train.py:380-395, TensorTrain.multiply). */
int ttx_hadamard_core_f64(ttx_stream_t stream, const double* A, int64_t ra, int64_t ni,
                          int64_t nj, int64_t rb, const double* B, int64_t sa, int64_t sb,
                          double* out);

/* The tiling of ttx_gemm_f64 (no counterpart in the reference: a tuning knob of
 * this library).  A contraction whose 64 x 64 tile grid (batch included) has
 * fewer than `tiles` tiles runs on 32 x 32 tiles, which give identical values.
 * Process-wide; default 512.  Set by ttnumpy_amd.autotune from timings. */
int ttx_gemm_set_small_tile_limit(int64_t tiles);
int64_t ttx_gemm_small_tile_limit(void);

/* out[(a m), i, j, (b n)] = sum_k A[a, i, k, b] * B[m, k, j, n] for cores A
 * (ra, ni, nk, rb) and B (sa, nk, nj, sb): one core of the matrix product of
 * two trains, TTM @ TTM or TT @ TT^T (ttnumpy/tensor_train.py:427-499, dot). */
int ttx_ttm_ttm_core_f64(ttx_stream_t stream, const double* A, int64_t ra, int64_t ni,
                         int64_t nk, int64_t rb, const double* B, int64_t sa, int64_t nj,
                         int64_t sb, double* out);

/* x[i] /= max_j |x[j]| when that maximum is positive: the scaling of each
 * random sketch of the randomized rounding (ttnumpy/utils.py:187-189). */
int ttx_scale_maxabs_f64(ttx_stream_t stream, int64_t n, double* x);

/* truncated_bond with delta = scale * nrm[0] read on the device (ttnumpy/utils.py:388-390,437-440). */
int ttx_truncated_bond_dev_f64(ttx_stream_t stream, const double* S, int64_t len,
                               const double* nrm, double scale, int64_t max_rank, int32_t* out);

/*
 * sigma-cutoff rank (ttnumpy/solvers/common.py:872-885): out[0] = max(#{S/S[0] > thr}, 1)
 * capped by max_rank (<= 0: no cap); out[1] = 1 when the cap bound it.
 */
int ttx_rank_by_threshold_f64(ttx_stream_t stream, const double* S, int64_t len, double thr,
                              int64_t max_rank, int32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* TTX_H_ */
