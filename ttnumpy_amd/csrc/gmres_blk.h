// Restarted GMRES by one block: the iteration of scipy.sparse.linalg.gmres
// (SciPy 1.15, _isolve/iterative.py; no preconditioner, atol = 0,
// restart = min(20, n), maxiter = 10 n), which the reference calls at
// ttnumpy/solvers/common.py:786.  Modified Gram-Schmidt Arnoldi, Givens
// rotations as LAPACK's dlartg makes them, the inner tolerance ptol adapted
// between cycles, and the true residual checked after each cycle.  The caller
// applies the operator: V[col] -> V[col+1] before each arnoldi step and
// x -> ax before each check.  The standalone kernels of krylov.hip and the
// engine's local solve run these same routines.
//
// work (doubles): V (m+1) x n | x | ax | r | H (m x (m+1), H[col][k]) | G (2m) | S (m+1)
// st (TTX_GMRES_STATE doubles): the scalars below.  Every routine starts and
// ends with a block barrier, so the state thread 0 writes is seen by the
// next routine in the same kernel as in the next kernel.
#pragma once

#include <float.h>

#include "ttx_common.h"

namespace ttx {

enum {
  G_DONE = 0, G_SKIP, G_INFO, G_COL, G_BREAK, G_PRESID, G_PTOL, G_PMAX, G_ATOL, G_BNRM,
  G_RNORM, G_ITER, G_MAXIT, G_INNER
};
constexpr int kGmresMaxRestart = 64;

struct GmresView {
  double *V, *x, *ax, *r, *H, *G, *S;
  __host__ __device__ GmresView(double* w, int64_t n, int64_t m)
      : V(w), x(w + (m + 1) * n), ax(w + (m + 2) * n), r(w + (m + 3) * n),
        H(w + (m + 4) * n), G(w + (m + 4) * n + m * (m + 1)),
        S(w + (m + 4) * n + m * (m + 1) + 2 * m) {}
};

__host__ __device__ inline int64_t gmres_work_doubles(int64_t n, int64_t m) {
  return (m + 4) * n + m * (m + 1) + 3 * m + 1;
}

// c, s, r with [c s; -s c] [f; g] = [r; 0], r carrying the sign of f, as
// LAPACK 3.10's dlartg computes them (scaled only outside sqrt(safmin)..sqrt(safmax/2))
__device__ inline void givens_lartg(double f, double g, double* c, double* s, double* r) {
  const double safmin = DBL_MIN, safmax = 1.0 / DBL_MIN;
  const double rtmin = sqrt(safmin), rtmax = sqrt(safmax / 2.0);
  const double f1 = fabs(f), g1 = fabs(g);
  if (g == 0.0) {
    *c = 1.0; *s = 0.0; *r = f;
  } else if (f == 0.0) {
    *c = 0.0; *s = copysign(1.0, g); *r = g1;
  } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
    const double d = sqrt(f * f + g * g);
    *c = f1 / d;
    *r = copysign(d, f);
    *s = g / *r;
  } else {
    const double u = fmin(safmax, fmax(safmin, fmax(f1, g1)));
    const double fs = f / u, gs = g / u;
    const double d = sqrt(fs * fs + gs * gs);
    *c = fabs(fs) / d;
    const double rr = copysign(d, f);
    *s = gs / rr;
    *r = rr * u;
  }
}

__device__ __forceinline__ double block_nrm2(int64_t n, const double* a, double* red) {
  double sq = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) sq += a[i] * a[i];
  return sqrt(block_sum(sq, red));
}

// start of a cycle from r: v0 = r / |r|, S = (|r|, 0, ...)
__device__ inline void gmres_start_cycle(int64_t n, int64_t m, GmresView& g, double rn, double* st) {
  const double inv = 1.0 / rn;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) g.V[i] = g.r[i] * inv;
  if (threadIdx.x == 0) {
    for (int64_t k = 0; k <= m; ++k) g.S[k] = 0.0;
    g.S[0] = rn;
    st[G_BREAK] = 0.0;
    st[G_SKIP] = 0.0;
    st[G_COL] = 0.0;
  }
}

// x = x0 (or 0), r = b - A x0 given ax0 = A x0 (or NULL); x0 and ax0 may be
// the work's own x and ax
__device__ inline void gmres_init_blk(int64_t n, int64_t m, const double* b, const double* x0,
                                      const double* ax0, double rtol, int64_t maxiter,
                                      double* work, double* st, double* red) {
  __syncthreads();
  GmresView g(work, n, m);
  const double bnrm2 = block_nrm2(n, b, red);
  const double atol = rtol * bnrm2 > 0.0 ? rtol * bnrm2 : 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < TTX_GMRES_STATE; ++i) st[i] = 0.0;
    st[G_ATOL] = atol;
    st[G_BNRM] = bnrm2;
    st[G_MAXIT] = (double)maxiter;
    st[G_PMAX] = 1.0;
    st[G_PTOL] = bnrm2 * fmin(1.0, bnrm2 > 0.0 ? atol / bnrm2 : 1.0);
  }
  if (bnrm2 == 0.0) {  // scipy returns b itself: the zero vector
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) g.x[i] = b[i];
    __syncthreads();
    if (threadIdx.x == 0) { st[G_DONE] = 1.0; st[G_SKIP] = 1.0; }
    __syncthreads();
    return;
  }
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double xi = x0 ? x0[i] : 0.0;
    const double ri = ax0 ? b[i] - ax0[i] : b[i];
    g.x[i] = xi;
    g.r[i] = ri;
  }
  const double rn = block_nrm2(n, g.r, red);
  __syncthreads();
  if (rn < atol) {
    if (threadIdx.x == 0) { st[G_DONE] = 1.0; st[G_SKIP] = 1.0; st[G_RNORM] = rn; }
    __syncthreads();
    return;
  }
  gmres_start_cycle(n, m, g, rn, st);
  __syncthreads();
}

// one Arnoldi step after the caller has written w = A V[col] into V[col+1]
__device__ inline void gmres_arnoldi_blk(int64_t n, int64_t m, int64_t col, double* work, double* st,
                                         double* red) {
  __syncthreads();
  if (st[G_SKIP] != 0.0) return;  // uniform across the block
  GmresView g(work, n, m);
  double* w = g.V + (col + 1) * n;
  double* h = g.H + col * (m + 1);
  const double h0 = block_nrm2(n, w, red);
  for (int64_t k = 0; k <= col; ++k) {
    const double* vk = g.V + k * n;
    double d = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d += vk[i] * w[i];
    const double t = block_sum(d, red);
    if (threadIdx.x == 0) h[k] = t;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) w[i] -= t * vk[i];
  }
  const double h1 = block_nrm2(n, w, red);
  const bool breakdown = h1 <= DBL_EPSILON * h0;
  if (!breakdown) {
    const double inv = 1.0 / h1;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) w[i] *= inv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    h[col + 1] = breakdown ? 0.0 : h1;
    for (int64_t k = 0; k < col; ++k) {
      const double c = g.G[2 * k], sn = g.G[2 * k + 1];
      const double n0 = h[k], n1 = h[k + 1];
      h[k] = c * n0 + sn * n1;
      h[k + 1] = -sn * n0 + c * n1;
    }
    double c, sn, mag;
    givens_lartg(h[col], h[col + 1], &c, &sn, &mag);
    g.G[2 * col] = c;
    g.G[2 * col + 1] = sn;
    h[col] = mag;
    h[col + 1] = 0.0;
    const double tmp = -sn * g.S[col];
    g.S[col] = c * g.S[col];
    g.S[col + 1] = tmp;
    const double presid = fabs(tmp);
    st[G_PRESID] = presid;
    st[G_INNER] += 1.0;
    st[G_COL] = (double)col;
    if (breakdown) st[G_BREAK] = 1.0;
    if (presid <= st[G_PTOL] || breakdown || col == m - 1) st[G_SKIP] = 1.0;
  }
  __syncthreads();
}

// end of a cycle: y from the triangular system, x += y V
__device__ inline void gmres_finish_blk(int64_t n, int64_t m, double* work, double* st) {
  __shared__ double y[kGmresMaxRestart];
  __syncthreads();
  if (st[G_DONE] != 0.0) return;
  GmresView g(work, n, m);
  const int64_t col = (int64_t)st[G_COL];
  if (threadIdx.x == 0) {
    if (g.H[col * (m + 1) + col] == 0.0) g.S[col] = 0.0;
    for (int64_t k = 0; k <= col; ++k) y[k] = g.S[k];
    for (int64_t k = col; k > 0; --k) {
      if (y[k] != 0.0) {
        y[k] /= g.H[k * (m + 1) + k];
        const double t = y[k];
        for (int64_t j = 0; j < k; ++j) y[j] -= t * g.H[k * (m + 1) + j];
      }
    }
    if (y[0] != 0.0) y[0] /= g.H[0];
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double acc = 0.0;
    for (int64_t k = 0; k <= col; ++k) acc += y[k] * g.V[k * n + i];
    g.x[i] += acc;
  }
  __syncthreads();
}

// after the caller has written ax = A x: r = b - ax and the outer decisions
__device__ inline void gmres_check_blk(int64_t n, int64_t m, const double* b, double* work,
                                       double* st, double* red) {
  __shared__ double s_next;
  __syncthreads();
  if (st[G_DONE] != 0.0) return;
  GmresView g(work, n, m);
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) g.r[i] = b[i] - g.ax[i];
  const double rnorm = block_nrm2(n, g.r, red);
  if (threadIdx.x == 0) {
    st[G_RNORM] = rnorm;
    st[G_ITER] += 1.0;
    const double atol = st[G_ATOL];
    double next = 0.0;
    if (rnorm <= atol) {
      st[G_DONE] = 1.0;
      st[G_INFO] = 0.0;
    } else if (st[G_BREAK] != 0.0) {
      st[G_DONE] = 1.0;
      st[G_INFO] = st[G_MAXIT];
    } else {
      if (st[G_PRESID] <= st[G_PTOL])
        st[G_PMAX] = fmax(DBL_EPSILON, 0.25 * st[G_PMAX]);
      else
        st[G_PMAX] = fmin(1.0, 1.5 * st[G_PMAX]);
      st[G_PTOL] = st[G_PRESID] * fmin(st[G_PMAX], atol / rnorm);
      if (st[G_ITER] >= st[G_MAXIT]) {
        st[G_DONE] = 1.0;
        st[G_INFO] = st[G_MAXIT];
      } else {
        next = 1.0;
      }
    }
    if (st[G_DONE] != 0.0) st[G_SKIP] = 1.0;
    s_next = next;
  }
  __syncthreads();
  if (s_next != 0.0) {
    const double rn = block_nrm2(n, g.r, red);
    gmres_start_cycle(n, m, g, rn, st);
  }
  __syncthreads();
}

}  // namespace ttx
