// Dense FP64 factorizations of core unfoldings and local systems:
//   * Householder QR (economic), one workgroup per matrix;
//   * thin SVD: QR of the tall side, then one-sided (Hestenes) Jacobi on the
//     square factor with a round-robin pair schedule, one workgroup;
//   * LU with partial pivoting for the dense local system: a one-workgroup
//     panel kernel (panel staged in LDS, pivoting, row swaps, triangular
//     solve of the U12 block row) alternating with the MFMA GEMM for the
//     trailing update, then a blocked back substitution;
//   * the rank rules of the truncations.
//
// Numerics, not bits, follow LAPACK: QR/SVD factors are unique up to column
// signs, which every consumer in the solver path is invariant to.

#include <float.h>

#include <mutex>

#include "block_ops.h"

namespace ttx {

// dynamic LDS of the global-workspace QR/SVD kernels (blocked QR panels)
constexpr size_t kQrLds = 150 * 1024;

struct QrArgs {
  int64_t m, n;
  const double* A;
  int64_t a_rs, a_cs;
  double* Q;
  int64_t q_rs, q_cs;
  double* R;
  int64_t r_rs, r_cs;
  double* work;
  const int* skip;  // when non-null and *skip == 0, Q and R are already formed: return
};

// IN_LDS: the working matrix, tau and Q live in dynamic LDS (small cores);
// otherwise in the caller's global workspace.
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void qr_kernel(QrArgs p) {
  extern __shared__ double smem[];
  __shared__ double red[16];
  if (p.skip != nullptr && *p.skip == 0) return;
  const int64_t m = p.m, n = p.n, k = m < n ? m : n;
  double* W = IN_LDS ? smem : p.work;
  double* tau = W + m * n;
  double* Qw = tau + k;
  for (int64_t e = threadIdx.x; e < m * n; e += blockDim.x) {
    const int64_t i = e % m, j = e / m;
    W[e] = p.A[i * p.a_rs + j * p.a_cs];
  }
  __syncthreads();
  // global workspace: LDS holds the blocked factorization's panels
  block_householder_qr(m, n, W, tau, Qw, red, p.Q != nullptr, IN_LDS ? nullptr : smem,
                       IN_LDS ? 0 : (int64_t)(kQrLds / sizeof(double)));
  __syncthreads();
  if (p.Q) {
    for (int64_t e = threadIdx.x; e < m * k; e += blockDim.x) {
      const int64_t i = e % m, c = e / m;
      p.Q[i * p.q_rs + c * p.q_cs] = Qw[e];
    }
  }
  for (int64_t e = threadIdx.x; e < k * n; e += blockDim.x) {
    const int64_t i = e % k, c = e / k;
    p.R[i * p.r_rs + c * p.r_cs] = (i <= c) ? W[i + c * m] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// CholeskyQR2 for tall matrices whose Q is wanted: G = X^T X, L L^T = G,
// Q1 = X L^{-T}; the same on Q1 (G2, L2, Q = Q1 L2^{-T}); R = L2^T L1^T.  The
// Gram and triangular products run on the multi-workgroup MFMA GEMM, so a
// 512 x 128 factorization spreads over the chip instead of one CU; the
// Cholesky factor and its inverse are formed by one workgroup in LDS.
// Two passes give orthogonality to rounding when kappa(X)^2 eps << 1
// (Yamamoto, Nakatsukasa, Yanagisawa, Fukaya 2015).  The factorization is
// refused (flag set) when a first-pass pivot falls below kCqrPivot times the
// largest diagonal of G (kappa(X) beyond ~1e6) or is not finite, or when a
// second-pass diagonal leaves [0.9, 1.1] (Q1 not close to orthonormal); the
// Householder kernel queued after it then recomputes Q and R from X.  The
// result equals Householder QR up to rounding and the signs of R's diagonal
// (positive here), which every consumer is invariant to.
// ---------------------------------------------------------------------------
constexpr int64_t kCqrMinCols = 64, kCqrMaxCols = 128, kCqrMinRows = 256;
constexpr double kCqrPivot = 1e-12;
constexpr int kCqrCholThreads = 1024;

static bool cqr_applies(int64_t m, int64_t n) {
  return n >= kCqrMinCols && n <= kCqrMaxCols && m >= kCqrMinRows && m >= 2 * n;
}

// Work layout of the CholeskyQR2 path: G, L1, L1inv, L2, L2inv (n x n each,
// row-major), Q1 (m x n row-major); the flag lives past the Householder
// region so that a fallback kernel reads it before anything overwrites it.
static int64_t cqr_work(int64_t m, int64_t n) { return 5 * n * n + m * n; }
static int64_t hh_work(int64_t m, int64_t n) {
  const int64_t k = m < n ? m : n;
  return m * n + k + m * k;
}

// Ordering of the LDS accesses of one wave between steps of a routine that
// only that wave runs (the LDS serves a wave's accesses in issue order; the
// fence keeps the compiler from moving loads and stores across the step).
__device__ __forceinline__ void wave_step() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

constexpr int kCqrNb = 16;  // block width of the Cholesky factor and its inverse

// Upper Cholesky of the b x b diagonal block at (j0, j0) of A (row stride ld)
// by the calling wave; false when a pivot is <= floor_piv or not finite.
__device__ bool wave_chol_block(double* A, int ld, int j0, int b, double floor_piv) {
  const int lane = threadIdx.x & 63;
  for (int j = 0; j < b; ++j) {
    double* rj = A + (j0 + j) * ld + j0;
    const double d = rj[j];
    if (!(d > floor_piv) || !isfinite(d)) return false;  // uniform in the wave
    const double l = sqrt(d), rl = 1.0 / l;
    for (int k = j + 1 + lane; k < b; k += 64) rj[k] *= rl;
    wave_step();
    if (lane == 0) rj[j] = l;
    const int w = b - j - 1;
    for (int e = lane; e < w * w; e += 64) {
      const int i = j + 1 + e / w, k = j + 1 + e % w;
      if (k >= i) A[(j0 + i) * ld + j0 + k] -= rj[i] * rj[k];
    }
    wave_step();
  }
  return true;
}

// In-place inverse of the b x b upper block at (j0, j0) by the calling wave
// (LAPACK dtrti2, upper: column j becomes -(1/U_jj) Inv[0:j,0:j] U[0:j, j]).
__device__ void wave_inv_block(double* A, int ld, int j0, int b) {
  const int lane = threadIdx.x & 63;
  for (int j = 0; j < b; ++j) {
    const double inv = 1.0 / A[(j0 + j) * ld + j0 + j];
    double acc = 0.0;
    if (lane < j)
      for (int s = lane; s < j; ++s) acc += A[(j0 + lane) * ld + j0 + s] * A[(j0 + s) * ld + j0 + j];
    wave_step();
    if (lane < j) A[(j0 + lane) * ld + j0 + j] = -inv * acc;
    if (lane == 0) A[(j0 + j) * ld + j0 + j] = inv;
    wave_step();
  }
}

// L L^T = G and Linv = L^{-1} for the row-major n x n G (its upper triangle
// is read), both written row-major with zeros above the diagonal.  pass 0
// sets *flag (0 accepted, 1 refused); pass 1 only sets it to 1 when the
// second-pass diagonal is off.  n <= kCqrMaxCols.
//
// LDS holds U = L^T (upper, row stride n + 1, so that rows and columns both
// spread over the banks).  Blocked by kCqrNb: wave 0 factors the diagonal
// block, all threads solve the block row U12 = D^{-T} A12 (one thread per
// column) and take the rank-kCqrNb update of the trailing upper triangle.
// U^{-1} is then formed top-down in the same storage: T = Inv11 A12 by all
// threads while wave 0 inverts the diagonal block, then A12 = -T D^{-1}; and
// L^{-1} = (U^{-1})^T.  Five barriers per block.
__global__ __launch_bounds__(1024) void cqr_chol_kernel(int64_t n64, const double* G, double* L,
                                                        double* Linv, int* flag, int pass) {
  extern __shared__ double A[];  // n x (n + 1), then T (n x kCqrNb)
  __shared__ int s_bad;
  __shared__ double s_gmax;
  const int n = (int)n64, ld = n + 1;
  double* T = A + n * ld;
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int wave = tid >> 6;
  const int tx = tid & 31, ty = tid >> 5, nty = nt >> 5;
  if (tid == 0) s_bad = 0;
  for (int e = tid; e < n * n; e += nt) {
    const int i = e / n, k = e - i * n;
    A[i * ld + k] = k >= i ? G[e] : 0.0;
  }
  __syncthreads();
  if (wave == 0) {  // largest diagonal entry of G
    double gm = 0.0;
    for (int i = tid; i < n; i += 64) gm = fmax(gm, A[i * ld + i]);
    gm = wave_max(gm);
    if (tid == 0) s_gmax = gm;
  }
  __syncthreads();
  const double floor_piv = pass == 0 ? kCqrPivot * s_gmax : 0.0;
  for (int jb = 0; jb < n; jb += kCqrNb) {
    const int b = n - jb < kCqrNb ? n - jb : kCqrNb;
    if (wave == 0 && !wave_chol_block(A, ld, jb, b, floor_piv) && tid == 0) s_bad = 1;
    __syncthreads();
    if (s_bad) break;  // uniform after the barrier
    const int c0 = jb + b;
    // U12 = D^{-T} A12: forward substitution down the block rows, one column per thread
    for (int c = c0 + tid; c < n; c += nt) {
      double x[kCqrNb];
#pragma unroll
      for (int r = 0; r < kCqrNb; ++r) {
        if (r < b) {
          double v = A[(jb + r) * ld + c];
#pragma unroll
          for (int q = 0; q < kCqrNb; ++q)
            if (q < r) v -= A[(jb + q) * ld + jb + r] * x[q];
          x[r] = v / A[(jb + r) * ld + jb + r];
          A[(jb + r) * ld + c] = x[r];
        }
      }
    }
    __syncthreads();
    // trailing upper triangle: A[i][k] -= sum_r U12[r][i] U12[r][k]
    for (int i = c0 + ty; i < n; i += nty) {
      for (int k = i + tx; k < n; k += 32) {
        double acc = 0.0;
        for (int r = 0; r < b; ++r) acc += A[(jb + r) * ld + i] * A[(jb + r) * ld + k];
        A[i * ld + k] -= acc;
      }
    }
    __syncthreads();
  }
  if (pass == 1 && !s_bad && tid == 0) {
    for (int j = 0; j < n; ++j)
      if (fabs(A[j * ld + j] - 1.0) > 0.1) s_bad = 1;
  }
  for (int e = tid; e < n * n; e += nt) {
    const int i = e / n, k = e - i * n;
    L[e] = k <= i ? A[k * ld + i] : 0.0;
  }
  __syncthreads();
  if (!s_bad) {
    for (int jb = 0; jb < n; jb += kCqrNb) {
      const int b = n - jb < kCqrNb ? n - jb : kCqrNb;
      // T = Inv11 A[0:jb, jb:jb+b] (Inv11 upper, already formed)
      for (int e = tid; e < jb * b; e += nt) {
        const int r = e / b, c = e - r * b;
        double acc = 0.0;
        for (int q = r; q < jb; ++q) acc += A[r * ld + q] * A[q * ld + jb + c];
        T[r * kCqrNb + c] = acc;
      }
      if (wave == 0) wave_inv_block(A, ld, jb, b);
      __syncthreads();
      // A[0:jb, jb:jb+b] = -T D^{-1} (D^{-1} upper)
      for (int e = tid; e < jb * b; e += nt) {
        const int r = e / b, c = e - r * b;
        double acc = 0.0;
        for (int q = 0; q <= c; ++q) acc += T[r * kCqrNb + q] * A[(jb + q) * ld + jb + c];
        A[r * ld + jb + c] = -acc;
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < n * n; e += nt) {
    const int i = e / n, k = e - i * n;
    Linv[e] = k <= i ? A[k * ld + i] : 0.0;
  }
  if (tid == 0) {
    if (pass == 0) *flag = s_bad;
    else if (s_bad) *flag = 1;
  }
}

// ---------------------------------------------------------------------------
// SVD
// ---------------------------------------------------------------------------
template <bool IN_LDS, bool ACCV = true>
__global__ __launch_bounds__(1024) void svd_kernel(SvdArgs p) {
  extern __shared__ double smem[];
  __shared__ double red[16];
  if (p.skip != nullptr && *p.skip == 0) return;
  block_svd<true, ACCV, !IN_LDS>(p, IN_LDS ? smem : p.work, red, IN_LDS ? nullptr : smem,
                        IN_LDS ? 0 : (int64_t)(kQrLds / sizeof(double)));
}

// *flag = 0 when max |E - I| <= tol for the n x n row-major E, else 1
__global__ __launch_bounds__(1024) void identity_check_kernel(int64_t n, const double* E,
                                                              double tol, int* flag) {
  __shared__ double red[16];
  double mx = 0.0;
  for (int64_t e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int64_t i = e / n, j = e - i * n;
    const double d = fabs(E[e] - (i == j ? 1.0 : 0.0));
    mx = (d > mx || d != d) ? d : mx;  // a NaN propagates
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(mx, off, 64);
    mx = (o > mx || o != o) ? o : mx;
  }
  if (lane == 0) red[wave] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = (red[w] > t || red[w] != red[w]) ? red[w] : t;
    *flag = (t <= tol) ? 0 : 1;
  }
}

// ---------------------------------------------------------------------------
// LU with partial pivoting and the solve of [M | b] (row-major) by one block:
// the same blocked routine the solver engine runs per request.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void chol_solve_kernel(int64_t n, int64_t ld, double* W,
                                                         double* u, int32_t* info) {
  extern __shared__ double lds[];
  if (threadIdx.x == 0) info[0] = TTX_OK;
  __syncthreads();
  block_chol_solve<256>(n, ld, W, u, lds, info);
}

__global__ __launch_bounds__(512) void lu_solve_kernel(int64_t n, int64_t ld, double* W, double* u,
                                                       int32_t* info, int64_t lds_bytes) {
  extern __shared__ double lds[];
  if (threadIdx.x == 0) info[0] = TTX_OK;
  block_lu_solve<512>(n, ld, W, u, lds, lds_bytes, info);
}


// ---------------------------------------------------------------------------
// rank rules
// ---------------------------------------------------------------------------
__global__ void truncated_bond_kernel(const double* S, int64_t len, double delta, int64_t max_rank,
                                      int32_t* out) {
  if (threadIdx.x != 0) return;
  // zero the rounding tail, then count tail norms above delta
  const double cutoff = kEps * (double)len * (len > 0 ? S[0] : 0.0);
  double tail = 0.0;
  int64_t count = 0;
  const double d2 = delta * delta;
  for (int64_t i = len - 1; i >= 0; --i) {
    const double v = S[i] <= cutoff ? 0.0 : S[i];
    tail += v * v;
    if (tail > d2) ++count;
  }
  int64_t bond = count > 1 ? count : 1;
  if (max_rank > 0 && bond > max_rank) bond = max_rank;
  out[0] = (int32_t)bond;
}

__global__ __launch_bounds__(1024) void rank_by_residual_kernel(int64_t N, int64_t limit,
                                                               const double* S,
                                                               const double* contrib,
                                                               const double* b, const double* exact,
                                                               double damp, double half_tol,
                                                               double* partial, int32_t* out) {
  __shared__ double red[16];
  const int tid = threadIdx.x;
  if (S[0] == 0.0) {
    if (tid == 0) { out[0] = 1; out[1] = 0; }
    return;
  }
  double s = 0.0;
  for (int64_t i = tid; i < N; i += blockDim.x) {
    const double v = b[i];
    partial[i] = -v;
    s += v * v;
  }
  const double bnorm = sqrt(block_sum(s, red));
  const double a = damp * exact[0];
  const double f = half_tol * bnorm;
  const double thr = a > f ? a : f;
  for (int64_t r = 0; r < limit; ++r) {
    const double* cr = contrib + r * N;
    double q = 0.0;
    for (int64_t i = tid; i < N; i += blockDim.x) {
      const double v = partial[i] + cr[i];
      partial[i] = v;
      q += v * v;
    }
    const double nrm = sqrt(block_sum(q, red));
    if (nrm <= thr) {
      if (tid == 0) { out[0] = (int32_t)(r + 1); out[1] = 0; }
      return;
    }
  }
  if (tid == 0) { out[0] = (int32_t)limit; out[1] = 1; }
}

__global__ __launch_bounds__(256) void svd_terms_kernel(int64_t p, int64_t q, int64_t limit,
                                                       const double* U, int64_t u_rs, int64_t u_cs,
                                                       const double* S, const double* Vt,
                                                       int64_t v_rs, int64_t v_cs, double* T) {
  const int64_t N = p * q;
  const int64_t total = N * limit;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t row = e % N, r = e / N;
    const int64_t i = row / q, j = row % q;
    T[e] = U[i * u_rs + r * u_cs] * S[r] * Vt[r * v_rs + j * v_cs];
  }
}

}  // namespace ttx

using namespace ttx;

// ---------------------------------------------------------------------------
// QR of a large matrix across the chip: LAPACK's blocked Householder scheme
// (geqrf with larft/larfb) with kQrbNb-column panels.  One workgroup factors
// a panel (columns in global memory, L2-resident) and forms V and the
// compact-WY factor T; the trailing columns take C -= V (T^T (V^T C)) as three
// MFMA GEMMs over the whole chip.  Q, when wanted, is accumulated as dorgqr
// does: from [I; 0], the panels in reverse order, X[j0:, j0:] -= V (T (V^T X)).
// Work: W (m x n, column-major) | tau (k) | V | T | Y (nb x n) | Y2 (nb x n),
// V and T holding one panel (nb columns) without Q and every panel with it:
// qrb_work(m, n) doubles.
// ---------------------------------------------------------------------------
constexpr int64_t kQrbNb = 16;
constexpr int kQrbLog = 4;  // log2(kQrbNb)
constexpr int kQrbThreads = 1024;
constexpr int64_t kQrbMinK = 64;
constexpr int64_t kQrbMinElems = int64_t(1) << 17;

static bool qrb_applies(int64_t m, int64_t n) {
  const int64_t k = m < n ? m : n;
  return k >= kQrbMinK && m * n >= kQrbMinElems;
}
static int64_t qrb_work(int64_t m, int64_t n) {
  const int64_t k = m < n ? m : n, npan = (k + kQrbNb - 1) / kQrbNb;
  return m * n + k + m * (k > kQrbNb ? k : kQrbNb) + npan * kQrbNb * kQrbNb + 2 * kQrbNb * n;
}

// Sum over the wave of kQrbNb values per lane: recursive halving over the
// lane bits 5 .. 6-kQrbLog, so lane l ends with the total of entry
// (l >> (6 - kQrbLog)) & (kQrbNb - 1) (the lanes that differ only in the low
// bits hold the same bits).  Fixed order, so deterministic.
__device__ __forceinline__ double wave_sum_nb(double (&a)[kQrbNb], int lane) {
#pragma unroll
  for (int s = 0; s < kQrbLog; ++s) {
    const int half = (int)kQrbNb >> (s + 1), off = 32 >> s;
    const bool hi = (lane & off) != 0;
#pragma unroll
    for (int j = 0; j < half; ++j) {
      const double keep = hi ? a[half + j] : a[j];
      const double send = hi ? a[j] : a[half + j];
      a[j] = keep + __shfl_xor(send, off, kWave);
    }
  }
  double v = a[0];
#pragma unroll
  for (int off = (1 << (6 - kQrbLog)) >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// One panel: columns j0 .. j0+jb-1 of W, rows j0 .. m-1 (mp rows).  Column c
// is reduced by the reflector H = I - tau v v^T of LAPACK's dlarfg (v[c] = 1,
// beta = -sign(alpha) ||x||); one pass over the rows then forms, for every
// other panel column cc, the product v_c^T (column cc): the trailing columns'
// (cc > c) give the update w = tau v^T P, the earlier reflectors' (cc < c)
// give G[cc][c] = v_cc^T v_c, from which T's column c follows as in dlarft
// (forward, columnwise): T[0:c, c] = -tau T[0:c, 0:c] G[0:c, c], T[c][c] = tau.
// V (mp x jb, column-major with leading dimension ldv, unit diagonal, zeros
// above) is written as the reflectors form; R stays in W's upper triangle.
__global__ __launch_bounds__(kQrbThreads) void qrb_panel_kernel(int64_t m, int64_t j0, int64_t jb,
                                                               double* W, double* tau, double* V,
                                                               int64_t ldv, double* T) {
  __shared__ double red[kQrbThreads / 64];
  __shared__ double part[kQrbThreads / 64][kQrbNb];
  __shared__ double sG[kQrbNb][kQrbNb];
  __shared__ double sT[kQrbNb][kQrbNb];
  __shared__ double sW[kQrbNb];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwave = kQrbThreads / 64;
  const int64_t mp = m - j0;
  double* P = W + j0 * m + j0;  // panel column c at P + c * m, rows 0 .. mp-1
  double sq = 0.0;
  for (int64_t i = 1 + tid; i < mp; i += kQrbThreads) sq += P[i] * P[i];
  sq = block_sum(sq, red);
  for (int64_t c = 0; c < jb; ++c) {
    double* col = P + c * m;
    const double alpha = col[c];
    double t = 0.0, beta = alpha, scal = 1.0;
    if (sq != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sq), alpha);
      scal = 1.0 / (alpha - beta);
      t = (beta - alpha) / beta;
    }
    double* vc = V + c * ldv;
    if (tid < c) vc[tid] = 0.0;
    double d[kQrbNb];
#pragma unroll
    for (int cc = 0; cc < kQrbNb; ++cc) d[cc] = 0.0;
    // all kQrbNb loads of a row are issued before any is used: columns past the
    // panel read its last column (the partial is never used), column c reads
    // itself (likewise)
    for (int64_t i = c + tid; i < mp; i += kQrbThreads) {
      double x[kQrbNb];
#pragma unroll
      for (int cc = 0; cc < kQrbNb; ++cc) {
        const int64_t q = cc < jb ? cc : jb - 1;
        x[cc] = cc < c ? V[q * ldv + i] : P[q * m + i];
      }
      const double v = i == c ? 1.0 : col[i] * scal;
      vc[i] = v;
#pragma unroll
      for (int cc = 0; cc < kQrbNb; ++cc) d[cc] += v * x[cc];
    }
    const double tot = wave_sum_nb(d, lane);
    if ((lane & ((1 << (6 - kQrbLog)) - 1)) == 0) part[wave][(lane >> (6 - kQrbLog)) & (kQrbNb - 1)] = tot;
    __syncthreads();
    if (tid < jb) {
      double s = 0.0;
      for (int w = 0; w < nwave; ++w) s += part[w][tid];
      if (tid < c) sG[tid][c] = s;
      sW[tid] = tid > c ? t * s : 0.0;
    }
    __syncthreads();
    if (tid < jb) {
      double v = 0.0;
      if (tid < c) {
        double acc = 0.0;
        for (int64_t l = tid; l < c; ++l) acc += sT[tid][l] * sG[l][c];
        v = -t * acc;
      } else if (tid == c) {
        v = t;
      }
      sT[tid][c] = v;
    }
    if (tid == 0) {
      tau[j0 + c] = t;
      col[c] = beta;
    }
    double w[kQrbNb];
#pragma unroll
    for (int cc = 0; cc < kQrbNb; ++cc) w[cc] = sW[cc];
    double sqn = 0.0;
    if (t != 0.0) {
      for (int64_t i = c + tid; i < mp; i += kQrbThreads) {
        double x[kQrbNb];
#pragma unroll
        for (int cc = 0; cc < kQrbNb; ++cc) x[cc] = P[(cc < jb ? cc : jb - 1) * m + i];
        const double v = vc[i];
#pragma unroll
        for (int cc = 0; cc < kQrbNb; ++cc) {
          if (cc > c && cc < jb) {
            const double nv = x[cc] - w[cc] * v;
            P[cc * m + i] = nv;
            if (cc == c + 1 && i > c + 1) sqn += nv * nv;
          }
        }
      }
    } else if (c + 1 < jb) {
      const double* nx = P + (c + 1) * m;
      for (int64_t i = c + 2 + tid; i < mp; i += kQrbThreads) sqn += nx[i] * nx[i];
    }
    sq = block_sum(sqn, red);  // also orders this column's writes before the next
  }
  for (int64_t e = tid; e < jb * jb; e += kQrbThreads) {
    const int64_t i = q32(e, jb), c = e - i * jb;
    T[e] = sT[i][c];
  }
}

__global__ __launch_bounds__(256) void qrb_extract_r_kernel(int64_t m, int64_t n, int64_t k,
                                                           const double* W, double* R,
                                                           int64_t r_rs, int64_t r_cs) {
  const int64_t total = k * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q32(e, n), j = e - i * n;
    R[i * r_rs + j * r_cs] = i <= j ? W[j * m + i] : 0.0;
  }
}

// Q = [I_k; 0] (m x k, strides q_rs, q_cs)
__global__ void qrb_eye_kernel(int64_t m, int64_t k, double* Q, int64_t q_rs, int64_t q_cs) {
  const int64_t total = m * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q32(e, k), c = e - i * k;
    Q[i * q_rs + c * q_cs] = i == c ? 1.0 : 0.0;
  }
}

static unsigned qrb_grid(int64_t elems) {
  const int64_t b = (elems + 255) / 256;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

static int qr_blocked(ttx_stream_t stream, int64_t m, int64_t n, const double* A, int64_t a_rs,
                      int64_t a_cs, double* Q, int64_t q_rs, int64_t q_cs, double* R, int64_t r_rs,
                      int64_t r_cs, double* work) {
  const int64_t k = m < n ? m : n, npan = (k + kQrbNb - 1) / kQrbNb;
  const bool form_q = Q != nullptr;
  double* W = work;
  double* tau = W + m * n;
  double* V = tau + k;
  double* T = V + m * (form_q && k > kQrbNb ? k : kQrbNb);
  double* Y = T + (form_q ? npan : 1) * kQrbNb * kQrbNb;
  double* Y2 = Y + kQrbNb * n;
  hipStream_t st = (hipStream_t)stream;
  {
    const int64_t dims[2] = {n, m}, ss[2] = {a_cs, a_rs}, ds[2] = {m, 1};
    const int rc = ttx_copy_f64(stream, 2, dims, A, ss, W, ds, 1.0);
    if (rc != TTX_OK) return rc;
  }
  for (int64_t j0 = 0; j0 < k; j0 += kQrbNb) {
    const int64_t jb = (k - j0) < kQrbNb ? (k - j0) : kQrbNb;
    const int64_t mp = m - j0, nt = n - j0 - jb;
    // this panel's V (rows j0.., leading dimension ldv) and T
    double* Vp = form_q ? V + j0 * m + j0 : V;
    const int64_t ldv = form_q ? m : mp;
    double* Tp = form_q ? T + (j0 / kQrbNb) * kQrbNb * kQrbNb : T;
    hipLaunchKernelGGL(qrb_panel_kernel, dim3(1), dim3(kQrbThreads), 0, st, m, j0, jb, W, tau, Vp,
                       ldv, Tp);
    int rc = launch_status();
    if (rc != TTX_OK) return rc;
    if (nt <= 0) continue;
    double* C = W + (j0 + jb) * m + j0;  // mp x nt, column-major (ld m)
    // Y = V^T C (jb x nt, row-major)
    rc = ttx_gemm_f64(stream, 1, jb, nt, mp, 1.0, Vp, ldv, 1, 0, C, 1, m, 0, 0.0, Y, nt, 1, 0);
    if (rc != TTX_OK) return rc;
    // Y2 = T^T Y
    rc = ttx_gemm_f64(stream, 1, jb, nt, jb, 1.0, Tp, 1, jb, 0, Y, nt, 1, 0, 0.0, Y2, nt, 1, 0);
    if (rc != TTX_OK) return rc;
    // C -= V Y2
    rc = ttx_gemm_f64(stream, 1, mp, nt, jb, -1.0, Vp, 1, ldv, 0, Y2, nt, 1, 0, 1.0, C, 1, m, 0);
    if (rc != TTX_OK) return rc;
  }
  hipLaunchKernelGGL(qrb_extract_r_kernel, dim3(qrb_grid(k * n)), dim3(256), 0, st, m, n, k, W, R,
                     r_rs, r_cs);
  int rc = launch_status();
  if (rc != TTX_OK || !form_q) return rc;
  hipLaunchKernelGGL(qrb_eye_kernel, dim3(qrb_grid(m * k)), dim3(256), 0, st, m, k, Q, q_rs, q_cs);
  if ((rc = launch_status()) != TTX_OK) return rc;
  for (int64_t p = npan - 1; p >= 0; --p) {
    const int64_t j0 = p * kQrbNb;
    const int64_t jb = (k - j0) < kQrbNb ? (k - j0) : kQrbNb;
    const int64_t mp = m - j0, nc = k - j0;
    const double* Vp = V + j0 * m + j0;
    const double* Tp = T + p * kQrbNb * kQrbNb;
    double* X = Q + j0 * q_rs + j0 * q_cs;  // mp x nc
    // Y = V^T X, Y2 = T Y, X -= V Y2
    if ((rc = ttx_gemm_f64(stream, 1, jb, nc, mp, 1.0, Vp, m, 1, 0, X, q_rs, q_cs, 0, 0.0, Y, nc, 1,
                           0)) != TTX_OK)
      return rc;
    if ((rc = ttx_gemm_f64(stream, 1, jb, nc, jb, 1.0, Tp, jb, 1, 0, Y, nc, 1, 0, 0.0, Y2, nc, 1,
                           0)) != TTX_OK)
      return rc;
    if ((rc = ttx_gemm_f64(stream, 1, mp, nc, jb, -1.0, Vp, 1, m, 0, Y2, nc, 1, 0, 1.0, X, q_rs,
                           q_cs, 0)) != TTX_OK)
      return rc;
  }
  return TTX_OK;
}

extern "C" int64_t ttx_qr_work_size(int64_t m, int64_t n) {
  int64_t h = hh_work(m, n);
  if (qrb_applies(m, n) && qrb_work(m, n) > h) h = qrb_work(m, n);
  if (!cqr_applies(m, n)) return h;
  const int64_t c = cqr_work(m, n);
  return (h > c ? h : c) + 2;  // + the flag
}

constexpr size_t kLdsLimit = 150 * 1024;

static int set_lds_attr(const void* fn) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsLimit) ==
                 hipSuccess
             ? TTX_OK
             : TTX_EHIP;
}

// CholeskyQR2 of X (m x n, strided) into Q and R, with the Householder
// fallback queued behind it; work: ttx_qr_work_size(m, n) doubles.
static int cqr_factor(ttx_stream_t stream, int64_t m, int64_t n, const double* A, int64_t a_rs,
                      int64_t a_cs, double* Q, int64_t q_rs, int64_t q_cs, double* R,
                      int64_t r_rs, int64_t r_cs, double* work) {
  static bool attr_c = false, attr_g = false;
  if (!attr_c) {
    if (set_lds_attr((const void*)cqr_chol_kernel) != TTX_OK) return TTX_EHIP;
    attr_c = true;
  }
  if (!attr_g) {
    if (set_lds_attr((const void*)qr_kernel<false>) != TTX_OK) return TTX_EHIP;
    attr_g = true;
  }
  const int64_t nn = n * n;
  double* G = work;
  double* L1 = G + nn;
  double* L1i = L1 + nn;
  double* L2 = L1i + nn;
  double* L2i = L2 + nn;
  double* Q1 = L2i + nn;
  const int64_t h = hh_work(m, n), c = cqr_work(m, n);
  int* flag = (int*)(work + (h > c ? h : c));
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = (size_t)(n * (n + 1) + n * kCqrNb) * sizeof(double);
  int rc;
  // pass 1: G = X^T X, L1, Q1 = X L1^{-T}
  if ((rc = ttx_gemm_f64(stream, 1, n, n, m, 1.0, A, a_cs, a_rs, 0, A, a_rs, a_cs, 0, 0.0, G, n, 1,
                         0)) != TTX_OK)
    return rc;
  hipLaunchKernelGGL(cqr_chol_kernel, dim3(1), dim3(kCqrCholThreads), lds, st, n, G, L1, L1i, flag, 0);
  if ((rc = ttx_gemm_f64(stream, 1, m, n, n, 1.0, A, a_rs, a_cs, 0, L1i, 1, n, 0, 0.0, Q1, n, 1,
                         0)) != TTX_OK)
    return rc;
  // pass 2: G = Q1^T Q1, L2, Q = Q1 L2^{-T}, R = L2^T L1^T
  if ((rc = ttx_gemm_f64(stream, 1, n, n, m, 1.0, Q1, 1, n, 0, Q1, n, 1, 0, 0.0, G, n, 1, 0)) !=
      TTX_OK)
    return rc;
  hipLaunchKernelGGL(cqr_chol_kernel, dim3(1), dim3(kCqrCholThreads), lds, st, n, G, L2, L2i, flag, 1);
  if ((rc = ttx_gemm_f64(stream, 1, m, n, n, 1.0, Q1, n, 1, 0, L2i, 1, n, 0, 0.0, Q, q_rs, q_cs,
                         0)) != TTX_OK)
    return rc;
  if ((rc = ttx_gemm_f64(stream, 1, n, n, n, 1.0, L2, 1, n, 0, L1, 1, n, 0, 0.0, R, r_rs, r_cs,
                         0)) != TTX_OK)
    return rc;
  // refused: Householder from X (returns at once when the flag is 0)
  QrArgs pf{m, n, A, a_rs, a_cs, Q, q_rs, q_cs, R, r_rs, r_cs, work, flag};
  hipLaunchKernelGGL(qr_kernel<false>, dim3(1), dim3(1024), kQrLds, st, pf);
  return launch_status();
}

extern "C" int ttx_qr_f64(ttx_stream_t stream, int64_t m, int64_t n, const double* A, int64_t a_rs,
                          int64_t a_cs, double* Q, int64_t q_rs, int64_t q_cs, double* R,
                          int64_t r_rs, int64_t r_cs, double* work) {
  if (m < 1 || n < 1) return TTX_EINVAL;
  static bool attr = false;
  if (!attr) {
    if (set_lds_attr((const void*)qr_kernel<true>) != TTX_OK) return TTX_EHIP;
    attr = true;
  }
  if (Q != nullptr && cqr_applies(m, n))
    return cqr_factor(stream, m, n, A, a_rs, a_cs, Q, q_rs, q_cs, R, r_rs, r_cs, work);
  // large matrices: blocked across the chip
  if (qrb_applies(m, n))
    return qr_blocked(stream, m, n, A, a_rs, a_cs, Q, q_rs, q_cs, R, r_rs, r_cs, work);
  QrArgs p{m, n, A, a_rs, a_cs, Q, q_rs, q_cs, R, r_rs, r_cs, work, nullptr};
  const size_t bytes = (size_t)hh_work(m, n) * sizeof(double);
  const int64_t k = m < n ? m : n;
  // k >= kQrBlockMin: the blocked factorization (panels in LDS), as in the engine
  if (bytes <= kLdsLimit && k < kQrBlockMin) {
    hipLaunchKernelGGL(qr_kernel<true>, dim3(1), dim3(1024), bytes, (hipStream_t)stream, p);
  } else {
    static bool attr_g = false;
    if (!attr_g) {
      if (set_lds_attr((const void*)qr_kernel<false>) != TTX_OK) return TTX_EHIP;
      attr_g = true;
    }
    hipLaunchKernelGGL(qr_kernel<false>, dim3(1), dim3(1024), kQrLds, (hipStream_t)stream, p);
  }
  return launch_status();
}

static int64_t svd_block_work(int64_t P, int64_t Q) { return 2 * P * Q + 2 * Q * Q + 3 * Q + 8; }

// V_r from R^{-1} is accepted when max |V_r^T V_r - I| <= kSvdCert * Q * eps
constexpr double kSvdCert = 16.0;

// ---------------------------------------------------------------------------
// SVD of factors whose Q x Q Jacobi matrix does not fit LDS (Q >= kSvdmMinQ):
// the one-sided Jacobi runs across the chip, one launch per round of the
// round-robin schedule and one workgroup per column pair, instead of on one
// workgroup whose rounds wait on global memory.  Preconditioned as the
// one-workgroup kernel's large case (Drmac-Veselic): X's columns by
// decreasing norm (X Pi = Qx R, Householder with Q), the rotations on the
// columns of R^T (R's rows, which row-major R already stores as columns),
// R^T V = Y, then U_X = Qx V and V_X = Pi Y Sigma^-1.  The host reads the
// rotation count once per sweep (a stream synchronization per sweep), and
// stops at the first sweep that rotates nothing or after kSvdmMaxSweeps.
// Work: Xp (P x Q) | Qx (P x Q) | R (Q x Q) | V (Q x Q) | Vs (Q x Q) | sig (Q)
// | the QR's work | ints: perm (Q), ord (Q), flags ((QQ-1) QQ/2), count, bad.
// ---------------------------------------------------------------------------
// the smallest Q whose Q x Q Jacobi matrix exceeds the one-workgroup kernel's LDS
constexpr int64_t kSvdmMinQ = 139;
constexpr int kSvdmMaxSweeps = 80;
constexpr int kSvdmThreads = 256;

static int64_t svdm_work(int64_t P, int64_t Q) {
  const int64_t QQ = Q + (Q & 1);
  const int64_t ints = 2 * Q + (QQ - 1) * (QQ / 2) + 2;
  return 2 * P * Q + 3 * Q * Q + Q + ttx_qr_work_size(P, Q) + (ints + 1) / 2 + 1;
}

// squared column norms of X (P x Q, strides x_rs, x_cs), the source column of
// each position by decreasing norm (ties by index; a non-finite norm sorts
// last and sets *bad)
__global__ __launch_bounds__(1024) void svdm_presort_kernel(int64_t P, int64_t Q, const double* A,
                                                           int64_t x_rs, int64_t x_cs,
                                                           double* keys, int* perm, int* bad) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  if (tid == 0) *bad = 0;
  __syncthreads();
  for (int64_t j = wave; j < Q; j += nwave) {
    double s = 0.0;
    for (int64_t r = lane; r < P; r += 64) {
      const double v = A[r * x_rs + j * x_cs];
      s += v * v;
    }
    s = wave_sum(s);
    if (lane == 0) {
      if (!isfinite(s)) *bad = 1;
      keys[j] = isfinite(s) ? s : -1.0;
    }
  }
  __syncthreads();
  for (int64_t j = tid; j < Q; j += blockDim.x) {
    const double sj = keys[j];
    int64_t pos = 0;
    for (int64_t l = 0; l < Q; ++l) {
      const double sl = keys[l];
      if (sl > sj || (sl == sj && l < j)) ++pos;
    }
    perm[pos] = (int)j;
  }
}

// Xp (P x Q row-major) = X[:, perm]; V = I (Q x Q)
__global__ void svdm_gather_kernel(int64_t P, int64_t Q, const double* A, int64_t x_rs,
                                   int64_t x_cs, const int* perm, double* Xp, double* V) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P * Q; e += stride) {
    const int64_t i = q32(e, Q), c = e - i * Q;
    Xp[e] = A[i * x_rs + (int64_t)perm[c] * x_cs];
  }
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < Q * Q; e += stride)
    V[e] = (e == q32(e, Q) * (Q + 1)) ? 1.0 : 0.0;
}

// One round: workgroup pk rotates its pair of columns of G and of V (both
// column-major, ld Q) and records whether it rotated in flags[pk].
__global__ __launch_bounds__(kSvdmThreads) void svdm_round_kernel(int64_t Q, int64_t QQ, int64_t round,
                                                                  double* G, double* V, double tol,
                                                                  int* flags) {
  __shared__ double red[3][kSvdmThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t pk = blockIdx.x;
  // round-robin: position 0 stays, the others rotate by one per round
  int64_t i = 0, j;
  if (pk != 0) {
    i = pk - 1 + round;
    if (i >= QQ - 1) i -= QQ - 1;
    i += 1;
  }
  j = QQ - 2 - pk + round;
  if (j >= QQ - 1) j -= QQ - 1;
  j += 1;
  if (i >= Q || j >= Q) {
    if (tid == 0) flags[pk] = 0;
    return;
  }
  if (i > j) { const int64_t t = i; i = j; j = t; }
  double* gi = G + i * Q;
  double* gj = G + j * Q;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int64_t r = tid; r < Q; r += kSvdmThreads) {
    const double x = gi[r], y = gj[r];
    a += x * x;
    b += y * y;
    c += x * y;
  }
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  if (lane == 0) {
    red[0][wave] = a;
    red[1][wave] = b;
    red[2][wave] = c;
  }
  __syncthreads();
  a = b = c = 0.0;
  for (int w = 0; w < kSvdmThreads / 64; ++w) {
    a += red[0][w];
    b += red[1][w];
    c += red[2][w];
  }
  double cs, sn;
  const bool rot = jac_params(a, b, c, tol, cs, sn);
  if (tid == 0) flags[pk] = rot ? 1 : 0;
  if (!rot) return;
  double* vi = V + i * Q;
  double* vj = V + j * Q;
  for (int64_t r = tid; r < Q; r += kSvdmThreads) {
    const double x = gi[r], y = gj[r], u = vi[r], w = vj[r];
    gi[r] = jac_rot_lo(cs, sn, x, y);
    gj[r] = jac_rot_hi(cs, sn, x, y);
    vi[r] = jac_rot_lo(cs, sn, u, w);
    vj[r] = jac_rot_hi(cs, sn, u, w);
  }
}

__global__ __launch_bounds__(1024) void svdm_count_kernel(int64_t n, const int* flags, int* count) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) s += (double)flags[e];
  s = block_sum(s, red);
  if (threadIdx.x == 0) *count = (int)s;
}

// sig[j] = ||G[:, j]||, ord = positions by decreasing sig (ties by index; a
// NaN sorts last, so that ord is always a permutation)
__global__ __launch_bounds__(1024) void svdm_sigma_kernel(int64_t Q, const double* G, double* sig,
                                                         int* ord) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  for (int64_t j = wave; j < Q; j += nwave) {
    const double* gj = G + j * Q;
    double s = 0.0;
    for (int64_t r = lane; r < Q; r += 64) s += gj[r] * gj[r];
    s = wave_sum(s);
    if (lane == 0) sig[j] = sqrt(s);
  }
  __syncthreads();
  for (int64_t j = tid; j < Q; j += blockDim.x) {
    const double sj = sig[j] == sig[j] ? sig[j] : -1.0;
    int64_t pos = 0;
    for (int64_t l = 0; l < Q; ++l) {
      const double sl = sig[l] == sig[l] ? sig[l] : -1.0;
      if (sl > sj || (sl == sj && l < j)) ++pos;
    }
    ord[pos] = (int)j;
  }
}

// Vs (Q x Q row-major) = V[:, ord]; V_X[perm[j], o] = G[j, ord[o]] / sig into
// out (row stride o_rs, column stride o_cs); S; the status
__global__ void svdm_finish_kernel(int64_t Q, const double* G, const double* V, const double* sig,
                                   const int* ord, const int* perm, double* Vs, double* out,
                                   int64_t o_rs, int64_t o_cs, double* S, const int* bad,
                                   int converged, int32_t* info) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < Q * Q; e += stride) {
    const int64_t o = q32(e, Q), j = e - o * Q;
    const int64_t src = ord[o];
    const double s = sig[src];
    Vs[j * Q + o] = V[j + src * Q];
    out[(int64_t)perm[j] * o_rs + o * o_cs] = s > 0.0 ? G[j + src * Q] / s : 0.0;
    if (j == 0) S[o] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    info[0] = *bad ? TTX_ENONFINITE : (converged ? TTX_OK : TTX_ENOCONV);
}

__global__ void svdm_status_kernel(int32_t* info, int32_t code) { info[0] = code; }

// Orthonormal completion of the singular vectors of exactly zero singular
// values (complete_zero_columns on U's columns and Vt's rows); returns after
// one read of S when none is zero.
__global__ __launch_bounds__(1024) void svd_complete_kernel(int64_t m, int64_t n, double* U,
                                                            int64_t u_rs, int64_t u_cs,
                                                            const double* S, double* Vt,
                                                            int64_t v_rs, int64_t v_cs) {
  __shared__ double red[16];
  const int64_t k = m < n ? m : n;
  complete_zero_columns(m, k, U, u_rs, u_cs, S, red);
  complete_zero_columns(n, k, Vt, v_cs, v_rs, S, red);
}

static int svd_multi_jacobi(ttx_stream_t stream, int64_t m, int64_t n, const double* A,
                            int64_t a_rs, int64_t a_cs, double* U, int64_t u_rs, int64_t u_cs,
                            double* S, double* Vt, int64_t v_rs, int64_t v_cs, double* work,
                            int32_t* info) {
  const bool wide = m < n;
  const int64_t P = wide ? n : m, Q = wide ? m : n;
  const int64_t x_rs = wide ? a_cs : a_rs, x_cs = wide ? a_rs : a_cs;
  const int64_t QQ = Q + (Q & 1), npairs = QQ / 2;
  double* Xp = work;
  double* Qx = Xp + P * Q;
  double* R = Qx + P * Q;
  double* V = R + Q * Q;
  double* Vs = V + Q * Q;
  double* sig = Vs + Q * Q;
  double* wq = sig + Q;
  int* perm = (int*)(wq + ttx_qr_work_size(P, Q));
  int* ord = perm + Q;
  int* flags = ord + Q;
  int* count = flags + (QQ - 1) * npairs;
  int* bad = count + 1;
  hipStream_t st = (hipStream_t)stream;
  int rc;
  hipLaunchKernelGGL(svdm_presort_kernel, dim3(1), dim3(1024), 0, st, P, Q, A, x_rs, x_cs, sig,
                     perm, bad);
  if ((rc = launch_status()) != TTX_OK) return rc;
  {
    int h = 0;
    if (hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return TTX_EHIP;
    if (h != 0) {  // NaN/inf input: reported, nothing factored
      hipLaunchKernelGGL(svdm_status_kernel, dim3(1), dim3(1), 0, st, info, (int32_t)TTX_ENONFINITE);
      return launch_status();
    }
  }
  const unsigned gb = (unsigned)((P * Q + 255) / 256 < 2048 ? (P * Q + 255) / 256 : 2048);
  hipLaunchKernelGGL(svdm_gather_kernel, dim3(gb), dim3(256), 0, st, P, Q, A, x_rs, x_cs, perm, Xp,
                     V);
  if ((rc = launch_status()) != TTX_OK) return rc;
  if ((rc = ttx_qr_f64(stream, P, Q, Xp, Q, 1, Qx, Q, 1, R, Q, 1, wq)) != TTX_OK) return rc;
  // the Jacobi matrix is R^T column-major: column c is row c of row-major R
  const double tol = (double)Q * kEps;
  bool converged = false;
  for (int sweep = 0; sweep < kSvdmMaxSweeps && !converged; ++sweep) {
    for (int64_t round = 0; round < QQ - 1; ++round) {
      hipLaunchKernelGGL(svdm_round_kernel, dim3((unsigned)npairs), dim3(kSvdmThreads), 0, st, Q,
                         QQ, round, R, V, tol, flags + round * npairs);
      if ((rc = launch_status()) != TTX_OK) return rc;
    }
    hipLaunchKernelGGL(svdm_count_kernel, dim3(1), dim3(1024), 0, st, (QQ - 1) * npairs, flags,
                       count);
    if ((rc = launch_status()) != TTX_OK) return rc;
    int h = 0;
    if (hipMemcpyAsync(&h, count, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return TTX_EHIP;
    converged = h == 0;
  }
  hipLaunchKernelGGL(svdm_sigma_kernel, dim3(1), dim3(1024), 0, st, Q, R, sig, ord);
  if ((rc = launch_status()) != TTX_OK) return rc;
  // tall: U = U_X, Vt[o][perm j] = V_X;  wide: U[perm j][o] = V_X, Vt = U_X^T
  double* out = wide ? U : Vt;
  const int64_t o_rs = wide ? u_rs : v_cs, o_cs = wide ? u_cs : v_rs;
  const unsigned fb = (unsigned)((Q * Q + 255) / 256 < 2048 ? (Q * Q + 255) / 256 : 2048);
  hipLaunchKernelGGL(svdm_finish_kernel, dim3(fb), dim3(256), 0, st, Q, R, V, sig, ord, perm, Vs,
                     out, o_rs, o_cs, S, bad, converged ? 1 : 0, info);
  if ((rc = launch_status()) != TTX_OK) return rc;
  // U_X = Qx Vs (P x Q)
  if (!wide)
    rc = ttx_gemm_f64(stream, 1, P, Q, Q, 1.0, Qx, Q, 1, 0, Vs, Q, 1, 0, 0.0, U, u_rs, u_cs, 0);
  else
    rc = ttx_gemm_f64(stream, 1, P, Q, Q, 1.0, Qx, Q, 1, 0, Vs, Q, 1, 0, 0.0, Vt, v_cs, v_rs, 0);
  if (rc != TTX_OK) return rc;
  // the finish kernel leaves zero singular vectors for zero singular values
  hipLaunchKernelGGL(svd_complete_kernel, dim3(1), dim3(1024), 0, st, m, n, U, u_rs, u_cs, S, Vt,
                     v_rs, v_cs);
  return launch_status();
}

extern "C" int64_t ttx_svd_work_size(int64_t m, int64_t n) {
  const int64_t P = m > n ? m : n, Q = m < n ? m : n;
  if (Q >= kSvdmMinQ) return svdm_work(P, Q);
  if (!cqr_applies(P, Q)) return svd_block_work(P, Q);
  // Qx, R, U_r, Vt_r, two Q x Q scratch, the QR's work, the square SVD's work, a flag
  return P * Q + 5 * Q * Q + ttx_qr_work_size(P, Q) + svd_block_work(Q, Q) + 2;
}

extern "C" int ttx_svd_f64(ttx_stream_t stream, int64_t m, int64_t n, const double* A,
                           int64_t a_rs, int64_t a_cs, double* U, int64_t u_rs, int64_t u_cs,
                           double* S, double* Vt, int64_t v_rs, int64_t v_cs, double* work,
                           int32_t* info) {
  if (m < 1 || n < 1) return TTX_EINVAL;
  static bool attr = false;
  if (!attr) {
    if (set_lds_attr((const void*)svd_kernel<true>) != TTX_OK) return TTX_EHIP;
    attr = true;
  }
  const int64_t P = m > n ? m : n, Qd = m < n ? m : n;
  if (Qd >= kSvdmMinQ)
    return svd_multi_jacobi(stream, m, n, A, a_rs, a_cs, U, u_rs, u_cs, S, Vt, v_rs, v_cs, work,
                            info);
  if (cqr_applies(P, Qd)) {
    // X = A (tall) or A^T (wide), P x Qd: X = Qx R by CholeskyQR2 across the
    // chip, then R = U_r S Vt_r, then tall: U = Qx U_r, Vt = Vt_r; wide:
    // U = Vt_r^T, Vt = (Qx U_r)^T.  The Jacobi kernel first runs without
    // accumulating rotations (its loop then touches only the square factor in
    // LDS) and V_r = R^{-1} U_r S is formed from the inverse Cholesky factors
    // the QR left behind, R^{-1} = L1^{-T} L2^{-T}; V_r is accepted when
    // max |V_r^T V_r - I| <= kSvdCert (its error is about kappa(R) eps),
    // otherwise the Jacobi kernel runs again accumulating the rotations.
    static bool attr_g = false, attr_f = false;
    if (!attr_g) {
      if (set_lds_attr((const void*)svd_kernel<false, true>) != TTX_OK) return TTX_EHIP;
      attr_g = true;
    }
    if (!attr_f) {
      if (set_lds_attr((const void*)svd_kernel<false, false>) != TTX_OK) return TTX_EHIP;
      attr_f = true;
    }
    const bool wide = m < n;
    const int64_t x_rs = wide ? a_cs : a_rs, x_cs = wide ? a_rs : a_cs;
    const int64_t qq = Qd * Qd;
    double* Qx = work;
    double* Rm = Qx + P * Qd;
    double* Ur = Rm + qq;
    double* Vtr = Ur + qq;
    double* T = Vtr + qq;
    double* E = T + qq;
    double* wq = E + qq;
    double* ws = wq + ttx_qr_work_size(P, Qd);
    int* vflag = (int*)(ws + svd_block_work(Qd, Qd));
    const double* L1i = wq + 2 * qq;  // the CholeskyQR2 work layout: G, L1, L1inv, L2, L2inv
    const double* L2i = wq + 4 * qq;
    hipStream_t st = (hipStream_t)stream;
    int rc = cqr_factor(stream, P, Qd, A, x_rs, x_cs, Qx, Qd, 1, Rm, Qd, 1, wq);
    if (rc != TTX_OK) return rc;
    SvdArgs pf{Qd, Qd, Rm, Qd, 1, Ur, Qd, 1, S, Vtr, Qd, 1, ws, info, nullptr};
    hipLaunchKernelGGL((svd_kernel<false, false>), dim3(1), dim3(1024), kQrLds, st, pf);
    if ((rc = launch_status()) != TTX_OK) return rc;
    // T = L2inv^T U_r; V_r = L1inv^T T written transposed into Vt_r; rows of Vt_r times S
    if ((rc = ttx_gemm_f64(stream, 1, Qd, Qd, Qd, 1.0, L2i, 1, Qd, 0, Ur, Qd, 1, 0, 0.0, T, Qd, 1,
                           0)) != TTX_OK)
      return rc;
    if ((rc = ttx_gemm_f64(stream, 1, Qd, Qd, Qd, 1.0, L1i, 1, Qd, 0, T, Qd, 1, 0, 0.0, Vtr, 1, Qd,
                           0)) != TTX_OK)
      return rc;
    if ((rc = ttx_scale_cols_f64(stream, Qd, Qd, Vtr, 1, Qd, S)) != TTX_OK) return rc;
    // E = V_r^T V_r = Vt_r Vt_r^T
    if ((rc = ttx_gemm_f64(stream, 1, Qd, Qd, Qd, 1.0, Vtr, Qd, 1, 0, Vtr, 1, Qd, 0, 0.0, E, Qd, 1,
                           0)) != TTX_OK)
      return rc;
    hipLaunchKernelGGL(identity_check_kernel, dim3(1), dim3(1024), 0, st, Qd, E,
                       kSvdCert * (double)Qd * kEps, vflag);
    SvdArgs pa{Qd, Qd, Rm, Qd, 1, Ur, Qd, 1, S, Vtr, Qd, 1, ws, info, vflag};
    hipLaunchKernelGGL((svd_kernel<false, true>), dim3(1), dim3(1024), kQrLds, st, pa);
    if ((rc = launch_status()) != TTX_OK) return rc;
    const int64_t dims[2] = {Qd, Qd};
    if (!wide) {
      if ((rc = ttx_gemm_f64(stream, 1, P, Qd, Qd, 1.0, Qx, Qd, 1, 0, Ur, Qd, 1, 0, 0.0, U, u_rs,
                             u_cs, 0)) != TTX_OK)
        return rc;
      const int64_t ss[2] = {Qd, 1}, ds[2] = {v_rs, v_cs};
      return ttx_copy_f64(stream, 2, dims, Vtr, ss, Vt, ds, 1.0);
    }
    const int64_t ss[2] = {1, Qd}, ds[2] = {u_rs, u_cs};  // U[i][o] = Vt_r[o][i]
    if ((rc = ttx_copy_f64(stream, 2, dims, Vtr, ss, U, ds, 1.0)) != TTX_OK) return rc;
    // Vt[o][r] = sum_l Qx[r][l] U_r[l][o]
    return ttx_gemm_f64(stream, 1, P, Qd, Qd, 1.0, Qx, Qd, 1, 0, Ur, Qd, 1, 0, 0.0, Vt, v_cs, v_rs,
                        0);
  }
  SvdArgs p{m, n, A, a_rs, a_cs, U, u_rs, u_cs, S, Vt, v_rs, v_cs, work, info};
  const size_t bytes = (size_t)svd_block_work(m > n ? m : n, m < n ? m : n) * sizeof(double);
  const int64_t k = m < n ? m : n;
  if (bytes <= kLdsLimit && k < kQrBlockMin) {
    hipLaunchKernelGGL(svd_kernel<true>, dim3(1), dim3(1024), bytes, (hipStream_t)stream, p);
  } else {
    static bool attr_g = false;
    if (!attr_g) {
      if (set_lds_attr((const void*)svd_kernel<false>) != TTX_OK) return TTX_EHIP;
      attr_g = true;
    }
    hipLaunchKernelGGL(svd_kernel<false>, dim3(1), dim3(1024), kQrLds, (hipStream_t)stream, p);
  }
  return launch_status();
}

extern "C" int64_t ttx_lu_work_size(int64_t n) { return (n + 1) / 2 + 1; /* reserved */ }

extern "C" int ttx_lu_solve_f64(ttx_stream_t stream, int64_t n, double* Mb, int64_t ld, double* u,
                                double* work, int32_t* info) {
  (void)work;
  if (n < 1 || ld < n + 1) return TTX_EINVAL;
  // the panel (nb >= 1 columns of n rows), the row permutation and the
  // staging buffers share the block's LDS
  if ((int64_t)kLdsLimit < 4 * n + 8 * (n > 4224 ? n : 4224)) return TTX_EINVAL;
  static bool attr_set = false;
  if (!attr_set) {
    if (set_lds_attr((const void*)lu_solve_kernel) != TTX_OK) return TTX_EHIP;
    attr_set = true;
  }
  hipLaunchKernelGGL(lu_solve_kernel, dim3(1), dim3(512), kLdsLimit, (hipStream_t)stream, n, ld, Mb,
                     u, info, (int64_t)kLdsLimit);
  return launch_status();
}

extern "C" int ttx_chol_solve_f64(ttx_stream_t stream, int64_t n, double* Mb, int64_t ld, double* u,
                                  int32_t* info) {
  if (n < 1 || ld < n + 1) return TTX_EINVAL;
  static bool attr_set = false;
  if (!attr_set) {
    if (set_lds_attr((const void*)chol_solve_kernel) != TTX_OK) return TTX_EHIP;
    attr_set = true;
  }
  hipLaunchKernelGGL(chol_solve_kernel, dim3(1), dim3(256), kCholLds * sizeof(double),
                     (hipStream_t)stream, n, ld, Mb, u, info);
  return launch_status();
}

extern "C" int ttx_truncated_bond_f64(ttx_stream_t stream, const double* S, int64_t len,
                                      double delta, int64_t max_rank, int32_t* out) {
  if (len < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(truncated_bond_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, S, len,
                     delta, max_rank, out);
  return launch_status();
}

extern "C" int ttx_rank_by_residual_f64(ttx_stream_t stream, int64_t N, int64_t limit,
                                        const double* S, const double* contrib,
                                        const double* b_loc, const double* exact, double damp,
                                        double half_trunc_tol, double* work, int32_t* out) {
  if (N < 1 || limit < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(rank_by_residual_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, N,
                     limit, S, contrib, b_loc, exact, damp, half_trunc_tol, work, out);
  return launch_status();
}

extern "C" int ttx_svd_terms_f64(ttx_stream_t stream, int64_t p, int64_t q, int64_t limit,
                                 const double* U, int64_t u_rs, int64_t u_cs, const double* S,
                                 const double* Vt, int64_t v_rs, int64_t v_cs, double* terms) {
  if (p < 1 || q < 1 || limit < 1) return TTX_EINVAL;
  int64_t total = p * q * limit;
  int64_t g = (total + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(svd_terms_kernel, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, p, q,
                     limit, U, u_rs, u_cs, S, Vt, v_rs, v_cs, terms);
  return launch_status();
}

// ---------------------------------------------------------------------------
// ttx_svd_f64_noalloc: ttx_svd_f64 with its work buffer in a file-scope device
// array (`scratch`) instead of a caller allocation, for callers that cannot
// allocate device memory.  Each device holds its own copy of the array from
// the load of this library's code object until the process ends; the host
// finds its address with hipGetSymbolAddress once per device (the array's
// name is only a host-side stand-in, never a device address) and passes that
// pointer on as ttx_svd_f64's `work`.  ttx_svd_f64 writes every part of its
// work it reads, so the array needs no clearing between calls.  Calls on one
// device share the array: two such calls in flight at once (on different
// streams, or from different threads) race.
constexpr int64_t kSvdScratchDoubles = int64_t(4) << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kSvdScratchDoubles];

namespace {
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchMaxDevices] = {};
}  // namespace

// Address of the current device's copy of `scratch` (looked up on the first
// call on that device, with that device current).
static int svd_scratch(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchMaxDevices) return TTX_EHIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
      return TTX_EHIP;
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return TTX_OK;
}

extern "C" int64_t ttx_svd_noalloc_max_work(void) { return kSvdScratchDoubles; }

extern "C" int ttx_svd_f64_noalloc(ttx_stream_t stream, int64_t m, int64_t n, const double* A,
                                   int64_t a_rs, int64_t a_cs, double* U, int64_t u_rs,
                                   int64_t u_cs, double* S, double* Vt, int64_t v_rs,
                                   int64_t v_cs, int32_t* info) {
  if (m < 1 || n < 1) return TTX_EINVAL;
  if (ttx_svd_work_size(m, n) > kSvdScratchDoubles) return TTX_ESCRATCH;
  double* w = nullptr;
  const int rc = svd_scratch(&w);
  if (rc != TTX_OK) return rc;
  return ttx_svd_f64(stream, m, n, A, a_rs, a_cs, U, u_rs, u_cs, S, Vt, v_rs, v_cs, w, info);
}
