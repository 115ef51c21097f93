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
// SVD
// ---------------------------------------------------------------------------
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void svd_kernel(SvdArgs p) {
  extern __shared__ double smem[];
  __shared__ double red[16];
  block_svd(p, IN_LDS ? smem : p.work, red, IN_LDS ? nullptr : smem,
            IN_LDS ? 0 : (int64_t)(kQrLds / sizeof(double)));
}

// ---------------------------------------------------------------------------
// LU with partial pivoting and the solve of [M | b] (row-major) by one block:
// the same blocked routine the solver engine runs per request.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void chol_solve_kernel(int64_t n, int64_t ld, double* W,
                                                         double* u, int32_t* info) {
  __shared__ double lds[4480];
  if (threadIdx.x == 0) info[0] = TTX_OK;
  __syncthreads();
  block_chol_solve<512>(n, ld, W, u, lds, info);
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
  hipLaunchKernelGGL(chol_solve_kernel, dim3(1), dim3(512), 0, (hipStream_t)stream, n, ld, Mb, u, info);
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
