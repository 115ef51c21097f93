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

#include "block_ops.h"

namespace ttx {

struct QrArgs {
  int64_t m, n;
  const double* A;
  int64_t a_rs, a_cs;
  double* Q;
  int64_t q_rs, q_cs;
  double* R;
  int64_t r_rs, r_cs;
  double* work;
};

// IN_LDS: the working matrix, tau and Q live in dynamic LDS (small cores);
// otherwise in the caller's global workspace.
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void qr_kernel(QrArgs p) {
  extern __shared__ double smem[];
  __shared__ double red[16];
  const int64_t m = p.m, n = p.n, k = m < n ? m : n;
  double* W = IN_LDS ? smem : p.work;
  double* tau = W + m * n;
  double* Qw = tau + k;
  for (int64_t e = threadIdx.x; e < m * n; e += blockDim.x) {
    const int64_t i = e % m, j = e / m;
    W[e] = p.A[i * p.a_rs + j * p.a_cs];
  }
  __syncthreads();
  block_householder_qr(m, n, W, tau, Qw, red, p.Q != nullptr);
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
// SVD
// ---------------------------------------------------------------------------
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void svd_kernel(SvdArgs p) {
  extern __shared__ double smem[];
  __shared__ double red[16];
  block_svd(p, IN_LDS ? smem : p.work, red);
}

// ---------------------------------------------------------------------------
// LU with partial pivoting on [M | b], column-major, leading dim n, n+1 cols.
// ---------------------------------------------------------------------------
struct PanelArgs {
  int64_t n, ld, k, nb;
  double* W;  // row-major [M | b], row stride ld >= n + 1
  int32_t* piv;
  int32_t* info;
};

// One panel of the right-looking LU: factor rows k.. of columns k..k+nb-1 in
// LDS with partial pivoting, apply the row swaps to every other column and
// solve the U12 block row.  Row-major storage makes the swaps and the
// triangular solve coalesced: consecutive threads own consecutive columns.
__global__ __launch_bounds__(512) void lu_panel_kernel(PanelArgs p) {
  extern __shared__ double P[];  // nb columns x mp rows, column-major
  __shared__ double rv[8];
  __shared__ int64_t ri[8];
  __shared__ int64_t s_piv[16];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  const int64_t n = p.n, ld = p.ld, k = p.k, nb = p.nb, ncols = n + 1;
  const int64_t mp = n - k;
  double* W = p.W;
  for (int64_t e = tid; e < nb * mp; e += blockDim.x) {
    const int64_t c = e % nb, r = e / nb;
    P[c * mp + r] = W[(k + r) * ld + (k + c)];
  }
  __syncthreads();
  for (int64_t j = 0; j < nb; ++j) {
    double* pj = P + j * mp;
    // pivot: first index of the largest magnitude (idamax semantics)
    double best = -1.0;
    int64_t bi = j;
    for (int64_t r = j + tid; r < mp; r += blockDim.x) {
      const double v = fabs(pj[r]);
      if (v > best) { best = v; bi = r; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off, 64);
      const int64_t oi = __shfl_xor(bi, off, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { rv[wave] = best; ri[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      double b0 = rv[0];
      int64_t i0 = ri[0];
      for (int w = 1; w < nwave; ++w)
        if (rv[w] > b0 || (rv[w] == b0 && ri[w] < i0)) { b0 = rv[w]; i0 = ri[w]; }
      s_piv[j] = i0;
      p.piv[k + j] = (int32_t)(k + i0);
      if (b0 == 0.0) p.info[0] = TTX_ESINGULAR;
    }
    __syncthreads();
    const int64_t pr = s_piv[j];
    if (pr != j) {
      for (int64_t c = tid; c < nb; c += blockDim.x) {
        double* pc = P + c * mp;
        const double t = pc[j];
        pc[j] = pc[pr];
        pc[pr] = t;
      }
    }
    __syncthreads();
    const double d = pj[j];
    if (d != 0.0) {
      for (int64_t r = j + 1 + tid; r < mp; r += blockDim.x) pj[r] /= d;
    }
    __syncthreads();
    for (int64_t c = j + 1; c < nb; ++c) {
      double* pc = P + c * mp;
      const double u = pc[j];
      for (int64_t r = j + 1 + tid; r < mp; r += blockDim.x) pc[r] -= pj[r] * u;
    }
    __syncthreads();
  }
  for (int64_t e = tid; e < nb * mp; e += blockDim.x) {
    const int64_t c = e % nb, r = e / nb;
    W[(k + r) * ld + (k + c)] = P[c * mp + r];
  }
  for (int64_t c = tid; c < ncols; c += blockDim.x) {
    if (c >= k && c < k + nb) continue;
    for (int64_t j = 0; j < nb; ++j) {
      const int64_t pr = k + s_piv[j];
      if (pr != k + j) {
        const double t = W[(k + j) * ld + c];
        W[(k + j) * ld + c] = W[pr * ld + c];
        W[pr * ld + c] = t;
      }
    }
    if (c >= k + nb) {
      for (int64_t i = 1; i < nb; ++i) {
        double s = W[(k + i) * ld + c];
        for (int64_t l = 0; l < i; ++l) s -= P[l * mp + i] * W[(k + l) * ld + c];
        W[(k + i) * ld + c] = s;
      }
    }
  }
}

// U x = y with y the last column (n) of the row-major factor; 64-row blocks,
// the diagonal block staged in LDS and solved by one wave with shuffles.
__global__ __launch_bounds__(1024) void lu_backsub_kernel(int64_t n, int64_t ld, const double* W,
                                                         double* u) {
  __shared__ double T[64][65];
  __shared__ double xb[64];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  for (int64_t i = tid; i < n; i += blockDim.x) u[i] = W[i * ld + n];
  __syncthreads();
  for (int64_t j1 = n; j1 > 0; j1 -= 64) {
    const int64_t j0 = j1 - 64 > 0 ? j1 - 64 : 0;
    const int64_t nbk = j1 - j0;
    for (int64_t e = tid; e < nbk * nbk; e += blockDim.x) {
      const int64_t r = e / nbk, c = e % nbk;
      T[r][c] = W[(j0 + r) * ld + j0 + c];
    }
    __syncthreads();
    if (wave == 0) {
      const int64_t row = j0 + lane;
      double yv = row < j1 ? u[row] : 0.0;
      for (int64_t j = j1 - 1; j >= j0; --j) {
        const int jj = (int)(j - j0);
        const double xj = __shfl(yv, jj, 64) / T[jj][jj];
        if (row == j) yv = xj;
        if (row < j) yv -= xj * T[lane][jj];
      }
      if (row < j1) { u[row] = yv; xb[lane] = yv; }
    }
    __syncthreads();
    for (int64_t r = tid; r < j0; r += blockDim.x) {
      double s = u[r];
      const double* wr = W + r * ld + j0;
      for (int64_t c = 0; c < nbk; ++c) s -= wr[c] * xb[c];
      u[r] = s;
    }
    __syncthreads();
  }
}

__global__ void lu_info_init_kernel(int32_t* info) { info[0] = TTX_OK; }

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

extern "C" int64_t ttx_qr_work_size(int64_t m, int64_t n) {
  const int64_t k = m < n ? m : n;
  return m * n + k + m * k;
}

constexpr size_t kLdsLimit = 150 * 1024;

static int set_lds_attr(const void* fn) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsLimit) ==
                 hipSuccess
             ? TTX_OK
             : TTX_EHIP;
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
  QrArgs p{m, n, A, a_rs, a_cs, Q, q_rs, q_cs, R, r_rs, r_cs, work};
  const size_t bytes = (size_t)ttx_qr_work_size(m, n) * sizeof(double);
  if (bytes <= kLdsLimit)
    hipLaunchKernelGGL(qr_kernel<true>, dim3(1), dim3(1024), bytes, (hipStream_t)stream, p);
  else
    hipLaunchKernelGGL(qr_kernel<false>, dim3(1), dim3(1024), 0, (hipStream_t)stream, p);
  return launch_status();
}

extern "C" int64_t ttx_svd_work_size(int64_t m, int64_t n) {
  const int64_t P = m > n ? m : n, Q = m < n ? m : n;
  return 2 * P * Q + 2 * Q * Q + 3 * Q + 8;
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
  SvdArgs p{m, n, A, a_rs, a_cs, U, u_rs, u_cs, S, Vt, v_rs, v_cs, work, info};
  const size_t bytes = (size_t)ttx_svd_work_size(m, n) * sizeof(double);
  if (bytes <= kLdsLimit)
    hipLaunchKernelGGL(svd_kernel<true>, dim3(1), dim3(1024), bytes, (hipStream_t)stream, p);
  else
    hipLaunchKernelGGL(svd_kernel<false>, dim3(1), dim3(1024), 0, (hipStream_t)stream, p);
  return launch_status();
}

static int64_t lu_block(int64_t n) {
  // the panel (n rows x nb cols) is staged in LDS
  int64_t nb = (int64_t)(kLdsLimit / (8 * (n > 0 ? n : 1)));
  if (nb > 16) nb = 16;
  return nb;
}

extern "C" int64_t ttx_lu_work_size(int64_t n) { return (n + 1) / 2 + 1; /* pivots as int32 */ }

extern "C" int ttx_lu_solve_f64(ttx_stream_t stream, int64_t n, double* Mb, int64_t ld, double* u,
                                double* work, int32_t* info) {
  if (n < 1 || ld < n + 1) return TTX_EINVAL;
  const int64_t nb0 = lu_block(n);
  if (nb0 < 1) return TTX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  static bool attr_set = false;
  if (!attr_set) {
    if (set_lds_attr((const void*)lu_panel_kernel) != TTX_OK) return TTX_EHIP;
    attr_set = true;
  }
  int32_t* piv = reinterpret_cast<int32_t*>(work);
  hipLaunchKernelGGL(lu_info_init_kernel, dim3(1), dim3(1), 0, st, info);
  for (int64_t k = 0; k < n; k += nb0) {
    const int64_t nb = (n - k) < nb0 ? (n - k) : nb0;
    PanelArgs pa{n, ld, k, nb, Mb, piv, info};
    const size_t shm = (size_t)(nb * (n - k)) * sizeof(double);
    hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(512), shm, st, pa);
    const int64_t rows = n - k - nb;
    const int64_t cols = n + 1 - k - nb;
    if (rows > 0 && cols > 0) {
      // A22 -= L21 * U12, all row-major with row stride ld
      int rc = ttx_gemm_f64(stream, 1, rows, cols, nb, -1.0, Mb + (k + nb) * ld + k, ld, 1, 0,
                            Mb + k * ld + (k + nb), ld, 1, 0, 1.0, Mb + (k + nb) * ld + (k + nb),
                            ld, 1, 0);
      if (rc != TTX_OK) return rc;
    }
  }
  hipLaunchKernelGGL(lu_backsub_kernel, dim3(1), dim3(1024), 0, st, n, ld, Mb, u);
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
