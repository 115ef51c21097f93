// Device-resident MINRES for the matrix-free local solve.
//
// The recurrences, the rotation updates and the stopping tests restate
// scipy.sparse.linalg.minres (SciPy 1.15, _isolve/minres.py, no
// preconditioner, shift 0), which the reference calls at
// ttnumpy/solvers/common.py:784.  The operator is applied between steps by
// the caller (the MFMA GEMM chain of the one-site apply); all scalar state
// stays on the device, so the host only polls `istop` every few iterations.
//
// vec layout (n doubles each): x, v, y, r1, r2, w, w1, w2.

#include <float.h>

#include "gmres_blk.h"
#include "ttx_common.h"

namespace ttx {

enum {
  S_ISTOP = 0, S_ITN, S_BETA1, S_BETA, S_OLDB, S_DBAR, S_EPSLN, S_PHIBAR, S_RHS1, S_RHS2,
  S_TNORM2, S_GMAX, S_GMIN, S_CS, S_SN, S_RTOL, S_MAXIT, S_ANORM, S_YNORM, S_ACOND, S_RNORM,
  S_BNORM
};

// istop value for the early returns of scipy (beta1 == 0 or ||b|| == 0)
constexpr double kIstopEarly = 100.0;

__global__ __launch_bounds__(1024) void minres_init_kernel(int64_t n, const double* b,
                                                          const double* x0, const double* ax0,
                                                          double rtol, int64_t maxiter,
                                                          double* vec, double* st) {
  __shared__ double red[16];
  double* x = vec;
  double* v = vec + n;
  double* r1 = vec + 3 * n;
  double* r2 = vec + 4 * n;
  double* w = vec + 5 * n;
  double* w1 = vec + 6 * n;
  double* w2 = vec + 7 * n;
  const int tid = threadIdx.x;
  double sb = 0.0, sr = 0.0;
  for (int64_t i = tid; i < n; i += blockDim.x) {
    const double xi = x0 ? x0[i] : 0.0;
    const double ri = ax0 ? b[i] - ax0[i] : b[i];
    x[i] = xi;
    r1[i] = ri;
    r2[i] = ri;
    w[i] = 0.0;
    w1[i] = 0.0;
    w2[i] = 0.0;
    sb += b[i] * b[i];
    sr += ri * ri;
  }
  const double beta1sq = block_sum(sr, red);
  const double bnorm = sqrt(block_sum(sb, red));
  double istop = 0.0;
  if (beta1sq == 0.0) {
    istop = kIstopEarly;  // x stays x0
  } else if (bnorm == 0.0) {
    istop = kIstopEarly;  // x = b
    for (int64_t i = tid; i < n; i += blockDim.x) x[i] = b[i];
  }
  const double beta1 = sqrt(beta1sq);
  if (istop == 0.0) {
    const double s = 1.0 / beta1;
    for (int64_t i = tid; i < n; i += blockDim.x) v[i] = s * r1[i];
  }
  if (tid == 0) {
    for (int i = 0; i < TTX_MINRES_STATE; ++i) st[i] = 0.0;
    st[S_ISTOP] = istop;
    st[S_ITN] = 0.0;
    st[S_BETA1] = beta1;
    st[S_BETA] = beta1;
    st[S_OLDB] = 0.0;
    st[S_DBAR] = 0.0;
    st[S_EPSLN] = 0.0;
    st[S_PHIBAR] = beta1;
    st[S_RHS1] = beta1;
    st[S_RHS2] = 0.0;
    st[S_TNORM2] = 0.0;
    st[S_GMAX] = 0.0;
    st[S_GMIN] = DBL_MAX;
    st[S_CS] = -1.0;
    st[S_SN] = 0.0;
    st[S_RTOL] = rtol;
    st[S_MAXIT] = (double)maxiter;
    st[S_BNORM] = bnorm;
  }
}

__global__ __launch_bounds__(1024) void minres_step_kernel(int64_t n, const double* y_in,
                                                          double* vec, double* st) {
  __shared__ double red[16];
  __shared__ double sc[8];
  if (st[S_ISTOP] != 0.0) return;  // uniform across the block
  const int tid = threadIdx.x;
  double* x = vec;
  double* v = vec + n;
  double* y = vec + 2 * n;
  double* r1 = vec + 3 * n;
  double* r2 = vec + 4 * n;
  double* w = vec + 5 * n;
  double* w1 = vec + 6 * n;
  double* w2 = vec + 7 * n;
  const double eps = DBL_EPSILON;
  const double itn = st[S_ITN] + 1.0;
  const double beta = st[S_BETA];
  const double oldb0 = st[S_OLDB];
  // y = A v - (beta/oldb) r1 ; alfa = <v, y>
  const double f1 = itn >= 2.0 ? beta / oldb0 : 0.0;
  double s = 0.0;
  for (int64_t i = tid; i < n; i += blockDim.x) {
    double yi = y_in[i];
    if (itn >= 2.0) yi = yi - f1 * r1[i];
    y[i] = yi;
    s += v[i] * yi;
  }
  const double alfa = block_sum(s, red);
  // y -= (alfa/beta) r2 ; r1 = r2 ; r2 = y ; beta' = ||r2||
  const double f2 = alfa / beta;
  s = 0.0;
  for (int64_t i = tid; i < n; i += blockDim.x) {
    const double yi = y[i] - f2 * r2[i];
    r1[i] = r2[i];
    r2[i] = yi;
    s += yi * yi;
  }
  const double betasq = block_sum(s, red);
  const double oldb = beta;
  const double nbeta = sqrt(betasq);
  // scalar recurrence (identical on every thread)
  double istop = st[S_ISTOP];
  const double beta1 = st[S_BETA1];
  double tnorm2 = st[S_TNORM2] + alfa * alfa + oldb * oldb + nbeta * nbeta;
  if (itn == 1.0 && nbeta / beta1 <= 10.0 * eps) istop = -1.0;
  const double cs0 = st[S_CS], sn0 = st[S_SN], dbar0 = st[S_DBAR];
  const double oldeps = st[S_EPSLN];
  const double delta = cs0 * dbar0 + sn0 * alfa;
  const double gbar = sn0 * dbar0 - cs0 * alfa;
  const double epsln = sn0 * nbeta;
  const double dbar = -cs0 * nbeta;
  const double root = sqrt(gbar * gbar + dbar * dbar);
  double gamma = sqrt(gbar * gbar + nbeta * nbeta);
  gamma = gamma > eps ? gamma : eps;
  const double cs = gbar / gamma;
  const double sn = nbeta / gamma;
  const double phibar0 = st[S_PHIBAR];
  const double phi = cs * phibar0;
  const double phibar = sn * phibar0;
  const double denom = 1.0 / gamma;
  // w1 = w2 ; w2 = w ; w = (v - oldeps w1 - delta w2) / gamma ; x += phi w
  s = 0.0;
  for (int64_t i = tid; i < n; i += blockDim.x) {
    const double ow2 = w2[i];
    const double ow = w[i];
    const double nw = (v[i] - oldeps * ow2 - delta * ow) * denom;
    w1[i] = ow2;
    w2[i] = ow;
    w[i] = nw;
    const double xi = x[i] + phi * nw;
    x[i] = xi;
    s += xi * xi;
  }
  const double ynorm = sqrt(block_sum(s, red));
  const double gmax = fmax(st[S_GMAX], gamma);
  const double gmin = fmin(st[S_GMIN], gamma);
  const double z = st[S_RHS1] / gamma;
  const double rhs1 = st[S_RHS2] - delta * z;
  const double rhs2 = -epsln * z;
  const double Anorm = sqrt(tnorm2);
  const double epsx = Anorm * ynorm * eps;
  const double rnorm = phibar;
  const double test1 = (ynorm == 0.0 || Anorm == 0.0) ? INFINITY : rnorm / (Anorm * ynorm);
  const double test2 = (Anorm == 0.0) ? INFINITY : root / Anorm;
  const double Acond = gmax / gmin;
  const double rtol = st[S_RTOL];
  if (istop == 0.0) {
    const double t1 = 1.0 + test1;
    const double t2 = 1.0 + test2;
    if (t2 <= 1.0) istop = 2.0;
    if (t1 <= 1.0) istop = 1.0;
    if (itn >= st[S_MAXIT]) istop = 6.0;
    if (Acond >= 0.1 / eps) istop = 4.0;
    if (epsx >= beta1) istop = 3.0;
    if (test2 <= rtol) istop = 2.0;
    if (test1 <= rtol) istop = 1.0;
  }
  if (istop == 0.0) {
    const double sv = 1.0 / nbeta;
    for (int64_t i = tid; i < n; i += blockDim.x) v[i] = sv * r2[i];
  }
  __syncthreads();  // every thread has read st[] before thread 0 rewrites it
  if (tid == 0) {
    st[S_ISTOP] = istop;
    st[S_ITN] = itn;
    st[S_BETA] = nbeta;
    st[S_OLDB] = oldb;
    st[S_DBAR] = dbar;
    st[S_EPSLN] = epsln;
    st[S_PHIBAR] = phibar;
    st[S_RHS1] = rhs1;
    st[S_RHS2] = rhs2;
    st[S_TNORM2] = tnorm2;
    st[S_GMAX] = gmax;
    st[S_GMIN] = gmin;
    st[S_CS] = cs;
    st[S_SN] = sn;
    st[S_ANORM] = Anorm;
    st[S_YNORM] = ynorm;
    st[S_ACOND] = Acond;
    st[S_RNORM] = rnorm;
  }
  (void)sc;
}

}  // namespace ttx

using namespace ttx;

extern "C" int ttx_minres_init_f64(ttx_stream_t stream, int64_t n, const double* b,
                                   const double* x0, const double* ax0, double rtol,
                                   int64_t maxiter, double* vec, double* st) {
  if (n < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(minres_init_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, b, x0,
                     ax0, rtol, maxiter, vec, st);
  return launch_status();
}

extern "C" int ttx_minres_step_f64(ttx_stream_t stream, int64_t n, const double* y_in, double* vec,
                                   double* st) {
  if (n < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(minres_step_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, y_in,
                     vec, st);
  return launch_status();
}

// ---------------------------------------------------------------------------
// Restarted GMRES (scipy.sparse.linalg.gmres, called at
// ttnumpy/solvers/common.py:786): one kernel per routine of gmres_blk.h, the
// operator applied by the caller between them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void gmres_init_kernel(int64_t n, int64_t m, const double* b,
                                                         const double* x0, const double* ax0,
                                                         double rtol, int64_t maxiter,
                                                         double* work, double* st) {
  __shared__ double red[16];
  gmres_init_blk(n, m, b, x0, ax0, rtol, maxiter, work, st, red);
}

__global__ __launch_bounds__(1024) void gmres_arnoldi_kernel(int64_t n, int64_t m, int64_t col,
                                                            double* work, double* st) {
  __shared__ double red[16];
  gmres_arnoldi_blk(n, m, col, work, st, red);
}

__global__ __launch_bounds__(1024) void gmres_finish_kernel(int64_t n, int64_t m, double* work,
                                                           double* st) {
  gmres_finish_blk(n, m, work, st);
}

__global__ __launch_bounds__(1024) void gmres_check_kernel(int64_t n, int64_t m, const double* b,
                                                          double* work, double* st) {
  __shared__ double red[16];
  gmres_check_blk(n, m, b, work, st, red);
}

extern "C" int64_t ttx_gmres_work_size(int64_t n, int64_t m) {
  if (n < 1 || m < 1) return 0;
  return (m + 4) * n + m * (m + 1) + 2 * m + (m + 1);
}

extern "C" int ttx_gmres_init_f64(ttx_stream_t stream, int64_t n, int64_t m, const double* b,
                                  const double* x0, const double* ax0, double rtol,
                                  int64_t maxiter, double* work, double* st) {
  if (n < 1 || m < 1 || m > kGmresMaxRestart || m > n || maxiter < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(gmres_init_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, m, b, x0,
                     ax0, rtol, maxiter, work, st);
  return launch_status();
}

extern "C" int ttx_gmres_arnoldi_f64(ttx_stream_t stream, int64_t n, int64_t m, int64_t col,
                                     double* work, double* st) {
  if (n < 1 || m < 1 || m > kGmresMaxRestart || col < 0 || col >= m) return TTX_EINVAL;
  hipLaunchKernelGGL(gmres_arnoldi_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, m, col,
                     work, st);
  return launch_status();
}

extern "C" int ttx_gmres_finish_f64(ttx_stream_t stream, int64_t n, int64_t m, double* work,
                                    double* st) {
  if (n < 1 || m < 1 || m > kGmresMaxRestart) return TTX_EINVAL;
  hipLaunchKernelGGL(gmres_finish_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, m, work,
                     st);
  return launch_status();
}

extern "C" int ttx_gmres_check_f64(ttx_stream_t stream, int64_t n, int64_t m, const double* b,
                                   double* work, double* st) {
  if (n < 1 || m < 1 || m > kGmresMaxRestart) return TTX_EINVAL;
  hipLaunchKernelGGL(gmres_check_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, m, b,
                     work, st);
  return launch_status();
}

// out[z, fl, i, fr] = alpha * sum psi[tl, a, fl] A[a, i, j, b] y[z, tl, j, tr] phi[tr, b, fr]
//                     + beta * out[z, fl, i, fr]
// as three strided GEMMs: s1 = y . phi, s2 = A . s1, out = psi^T . s2.
static int local_apply(ttx_stream_t stream, const double* psi, int64_t rl, int64_t Ra, int64_t fl,
                       const double* A, int64_t nr, int64_t nc, int64_t Rb, const double* phi,
                       int64_t rr, int64_t fr, const double* y, int64_t B, double alpha,
                       double beta, double* out, double* s1, double* s2, const double* skip) {
  if (rl < 1 || Ra < 1 || fl < 1 || nr < 1 || nc < 1 || Rb < 1 || rr < 1 || fr < 1 || B < 1)
    return TTX_EINVAL;
  int rc = gemm_f64_launch(stream, 1, B * rl * nc, Rb * fr, rr, 1.0, y, rr, 1, 0, phi, Rb * fr, 1,
                           0, 0.0, s1, Rb * fr, 1, 0, skip);
  if (rc != TTX_OK) return rc;
  rc = gemm_f64_launch(stream, B * rl, Ra * nr, fr, nc * Rb, 1.0, A, nc * Rb, 1, 0, s1, fr, 1,
                       nc * Rb * fr, 0.0, s2, fr, 1, Ra * nr * fr, skip);
  if (rc != TTX_OK) return rc;
  return gemm_f64_launch(stream, B, fl, nr * fr, rl * Ra, alpha, psi, 1, fl, 0, s2, nr * fr, 1,
                         rl * Ra * nr * fr, beta, out, nr * fr, 1, fl * nr * fr, skip);
}

extern "C" int ttx_local_apply_f64(ttx_stream_t stream, const double* psi, int64_t rl, int64_t Ra,
                                   int64_t fl, const double* A, int64_t nr, int64_t nc, int64_t Rb,
                                   const double* phi, int64_t rr, int64_t fr, const double* y,
                                   int64_t B, double alpha, double beta, double* out, double* s1,
                                   double* s2) {
  return local_apply(stream, psi, rl, Ra, fl, A, nr, nc, Rb, phi, rr, fr, y, B, alpha, beta, out,
                     s1, s2, nullptr);
}

extern "C" int ttx_minres_local_steps_f64(ttx_stream_t stream, const double* psi, int64_t rl,
                                          int64_t Ra, int64_t fl, const double* A, int64_t nr,
                                          int64_t nc, int64_t Rb, const double* phi, int64_t rr,
                                          int64_t fr, double* s1, double* s2, double* y_in,
                                          double* vec, double* st, int64_t count) {
  const int64_t n = fl * nr * fr;
  if (n < 1 || count < 0 || rl * nc * rr != n) return TTX_EINVAL;
  for (int64_t it = 0; it < count; ++it) {
    // once istop is set the matvec launches return at once, as the step does
    int rc = local_apply(stream, psi, rl, Ra, fl, A, nr, nc, Rb, phi, rr, fr, vec + n, 1, 1.0, 0.0,
                         y_in, s1, s2, st + S_ISTOP);
    if (rc != TTX_OK) return rc;
    rc = ttx_minres_step_f64(stream, n, y_in, vec, st);
    if (rc != TTX_OK) return rc;
  }
  return TTX_OK;
}

extern "C" int ttx_gmres_local_cycle_f64(ttx_stream_t stream, const double* psi, int64_t rl,
                                         int64_t Ra, int64_t fl, const double* A, int64_t nr,
                                         int64_t nc, int64_t Rb, const double* phi, int64_t rr,
                                         int64_t fr, double* s1, double* s2, int64_t m,
                                         const double* b, double* work, double* st) {
  const int64_t n = fl * nr * fr;
  if (n < 1 || rl * nc * rr != n || m < 1 || m > kGmresMaxRestart || m > n) return TTX_EINVAL;
  GmresView g(work, n, m);
  for (int64_t col = 0; col < m; ++col) {
    // the matvecs of a cycle whose inner loop has stopped return at once
    int rc = local_apply(stream, psi, rl, Ra, fl, A, nr, nc, Rb, phi, rr, fr, g.V + col * n, 1,
                         1.0, 0.0, g.V + (col + 1) * n, s1, s2, st + G_SKIP);
    if (rc != TTX_OK) return rc;
    rc = ttx_gmres_arnoldi_f64(stream, n, m, col, work, st);
    if (rc != TTX_OK) return rc;
  }
  int rc = ttx_gmres_finish_f64(stream, n, m, work, st);
  if (rc != TTX_OK) return rc;
  rc = local_apply(stream, psi, rl, Ra, fl, A, nr, nc, Rb, phi, rr, fr, g.x, 1, 1.0, 0.0, g.ax, s1,
                   s2, st + G_DONE);
  if (rc != TTX_OK) return rc;
  return ttx_gmres_check_f64(stream, n, m, b, work, st);
}
