// Tensor-train specific kernels that are not plain GEMMs.
//
//  * the explicit one-site local matrix, written column-major straight into
//    the [M | b] buffer of the LU solve (ttnumpy/solvers/one_site.py:49-75);
//  * one core of the TTM @ TT product, written into a block of the residual
//    train A x - b (ttnumpy/tensor_train.py:483-495, :300-318);
//  * column scaling U[:, j] *= s[j] (the U*S and S*Vt absorbs of the SVD
//    sweeps, ttnumpy/utils.py:400-404,449-453);
//  * truncated_bond with the budget read from a device norm (no host sync
//    for delta = error / sqrt(d-1) * ||core||, ttnumpy/utils.py:388-390).

#include <float.h>

#include "ttx_common.h"

namespace ttx {

struct LocalMatrixArgs {
  const double* psi;  // (rl, Ra, fl)
  const double* A;    // (Ra, nr, nc, Rb)
  const double* phi;  // (rr, Rb, fr)
  int64_t rl, Ra, fl, nr, nc, Rb, rr, fr;
  double* out;
  int64_t o_rs, o_cs;
};

__global__ __launch_bounds__(256) void local_matrix_kernel(LocalMatrixArgs p) {
  const int64_t rows = p.fl * p.nr * p.fr;
  const int64_t cols = p.rl * p.nc * p.rr;
  const int64_t total = rows * cols;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    // the index with the unit output stride varies fastest (coalesced stores)
    int64_t row, col;
    if (p.o_cs < p.o_rs) { col = e % cols; row = e / cols; } else { row = e % rows; col = e / rows; }
    const int64_t f_r = row % p.fr;
    const int64_t i = (row / p.fr) % p.nr;
    const int64_t f_l = row / (p.fr * p.nr);
    const int64_t t_r = col % p.rr;
    const int64_t j = (col / p.rr) % p.nc;
    const int64_t t_l = col / (p.rr * p.nc);
    p.out[row * p.o_rs + col * p.o_cs] =
        local_matrix_entry(p.psi + t_l * p.Ra * p.fl, p.fl, f_l, p.A + (i * p.nc + j) * p.Rb,
                           p.nr * p.nc * p.Rb, p.Ra, p.Rb, p.phi + t_r * p.Rb * p.fr, p.fr, f_r);
  }
}

struct TtmCoreArgs {
  const double* A;  // (Ra, nr, nc, Rb)
  const double* x;  // (rx, nc, rx2)
  int64_t Ra, nr, nc, Rb, rx, rx2;
  double* out;
  int64_t o_l, o_i, o_r;  // strides of the combined left index, row mode, right index
};

__global__ __launch_bounds__(256) void ttm_core_kernel(TtmCoreArgs p) {
  const int64_t L = p.Ra * p.rx, Rr = p.Rb * p.rx2;
  const int64_t total = L * p.nr * Rr;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t r = e % Rr;
    const int64_t i = (e / Rr) % p.nr;
    const int64_t l = e / (Rr * p.nr);
    const int64_t a = l / p.rx, m = l % p.rx;
    const int64_t b = r / p.rx2, nn = r % p.rx2;
    double s = 0.0;
    for (int64_t j = 0; j < p.nc; ++j)
      s += p.A[((a * p.nr + i) * p.nc + j) * p.Rb + b] * p.x[(m * p.nc + j) * p.rx2 + nn];
    p.out[l * p.o_l + i * p.o_i + r * p.o_r] = s;
  }
}

__global__ __launch_bounds__(256) void scale_cols_kernel(int64_t rows, int64_t cols, double* X,
                                                        int64_t rs, int64_t cs, const double* s) {
  const int64_t total = rows * cols;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t r = e % rows, c = e / rows;
    X[r * rs + c * cs] *= s[c];
  }
}

// out[(a, m), i, j, (b, n)] = A[a, i, j, b] B[m, i, j, n]: the core of the
// element-wise product of two trains (bonds multiply, modes shared)
__global__ __launch_bounds__(256) void hadamard_core_kernel(const double* A, int64_t ra, int64_t ni,
                                                           int64_t nj, int64_t rb, const double* B,
                                                           int64_t sa, int64_t sb, double* out) {
  const int64_t ob = rb * sb, total = ra * sa * ni * nj * ob;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    int64_t rem = e;
    const int64_t n = rem % sb; rem /= sb;
    const int64_t b = rem % rb; rem /= rb;
    const int64_t j = rem % nj; rem /= nj;
    const int64_t i = rem % ni; rem /= ni;
    const int64_t m = rem % sa;
    const int64_t a = rem / sa;
    out[e] = A[((a * ni + i) * nj + j) * rb + b] * B[((m * ni + i) * nj + j) * sb + n];
  }
}

// out[(a, m), i, j, (b, n)] = sum_k A[a, i, k, b] B[m, k, j, n]: one core of the
// matrix product of two trains (bonds multiply, the shared mode k summed)
__global__ __launch_bounds__(256) void ttm_ttm_core_kernel(const double* A, int64_t ra, int64_t ni,
                                                          int64_t nk, int64_t rb, const double* B,
                                                          int64_t sa, int64_t nj, int64_t sb,
                                                          double* out) {
  const int64_t total = ra * sa * ni * nj * rb * sb;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    int64_t rem = e;
    const int64_t n = rem % sb; rem /= sb;
    const int64_t b = rem % rb; rem /= rb;
    const int64_t j = rem % nj; rem /= nj;
    const int64_t i = rem % ni; rem /= ni;
    const int64_t m = rem % sa;
    const int64_t a = rem / sa;
    const double* pa = A + (a * ni + i) * nk * rb + b;       // A[a, i, k, b], k stride rb
    const double* pb = B + (m * nk * nj + j) * sb + n;       // B[m, k, j, n], k stride nj sb
    double acc = 0.0;
    for (int64_t k = 0; k < nk; ++k) acc = __fma_rn(pa[k * rb], pb[k * nj * sb], acc);
    out[e] = acc;
  }
}

// x /= max |x| when that is positive; one workgroup (the sketches it scales
// hold at most a few ten thousand entries)
__global__ __launch_bounds__(1024) void scale_maxabs_kernel(int64_t n, double* x) {
  __shared__ double red[16];
  __shared__ double s_top;
  double m = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(x[i]));
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = fmax(t, red[w]);
    s_top = t;
  }
  __syncthreads();
  const double top = s_top;
  if (top > 0.0)
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[i] /= top;
}

__global__ void truncated_bond_dev_kernel(const double* S, int64_t len, const double* nrm,
                                          double scale, int64_t max_rank, int32_t* out) {
  if (threadIdx.x != 0) return;
  const double delta = scale * nrm[0];
  const double cutoff = DBL_EPSILON * (double)len * S[0];
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

// rank = max(#{S[i] / S[0] > thr}, 1) capped by max_rank (ttnumpy/solvers/common.py:872-885);
// out[1] = 1 when the cap bound the rank.
__global__ void rank_by_threshold_kernel(const double* S, int64_t len, double thr, int64_t max_rank,
                                         int32_t* out) {
  if (threadIdx.x != 0) return;
  if (len == 0 || S[0] == 0.0) { out[0] = 1; out[1] = 0; return; }
  int64_t c = 0;
  for (int64_t i = 0; i < len; ++i) c += (S[i] / S[0] > thr) ? 1 : 0;
  int64_t r = c > 1 ? c : 1;
  int32_t capped = 0;
  if (max_rank > 0 && r > max_rank) { r = max_rank; capped = 1; }
  out[0] = (int32_t)r;
  out[1] = capped;
}

static inline unsigned grid_of(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (unsigned)g;
}

}  // namespace ttx

using namespace ttx;

extern "C" int ttx_local_matrix_f64(ttx_stream_t stream, const double* psi, int64_t rl, int64_t Ra,
                                    int64_t fl, const double* A, int64_t nr, int64_t nc,
                                    int64_t Rb, const double* phi, int64_t rr, int64_t fr,
                                    double* out, int64_t o_rs, int64_t o_cs) {
  if (rl < 1 || Ra < 1 || fl < 1 || nr < 1 || nc < 1 || Rb < 1 || rr < 1 || fr < 1) return TTX_EINVAL;
  LocalMatrixArgs p{psi, A, phi, rl, Ra, fl, nr, nc, Rb, rr, fr, out, o_rs, o_cs};
  const int64_t total = fl * nr * fr * rl * nc * rr;
  hipLaunchKernelGGL(local_matrix_kernel, dim3(grid_of(total)), dim3(256), 0, (hipStream_t)stream, p);
  return launch_status();
}

extern "C" int ttx_ttm_core_f64(ttx_stream_t stream, const double* A, int64_t Ra, int64_t nr,
                                int64_t nc, int64_t Rb, const double* x, int64_t rx, int64_t rx2,
                                double* out, int64_t o_l, int64_t o_i, int64_t o_r) {
  if (Ra < 1 || nr < 1 || nc < 1 || Rb < 1 || rx < 1 || rx2 < 1) return TTX_EINVAL;
  TtmCoreArgs p{A, x, Ra, nr, nc, Rb, rx, rx2, out, o_l, o_i, o_r};
  hipLaunchKernelGGL(ttm_core_kernel, dim3(grid_of(Ra * rx * nr * Rb * rx2)), dim3(256), 0,
                     (hipStream_t)stream, p);
  return launch_status();
}

extern "C" int ttx_scale_cols_f64(ttx_stream_t stream, int64_t rows, int64_t cols, double* X,
                                  int64_t rs, int64_t cs, const double* s) {
  if (rows < 0 || cols < 0) return TTX_EINVAL;
  if (rows == 0 || cols == 0) return TTX_OK;
  hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_of(rows * cols)), dim3(256), 0,
                     (hipStream_t)stream, rows, cols, X, rs, cs, s);
  return launch_status();
}

extern "C" int ttx_hadamard_core_f64(ttx_stream_t stream, const double* A, int64_t ra, int64_t ni,
                                     int64_t nj, int64_t rb, const double* B, int64_t sa,
                                     int64_t sb, double* out) {
  if (ra < 1 || ni < 1 || nj < 1 || rb < 1 || sa < 1 || sb < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(hadamard_core_kernel, dim3(grid_of(ra * sa * ni * nj * rb * sb)), dim3(256),
                     0, (hipStream_t)stream, A, ra, ni, nj, rb, B, sa, sb, out);
  return launch_status();
}

extern "C" int ttx_ttm_ttm_core_f64(ttx_stream_t stream, const double* A, int64_t ra,
                                    int64_t ni, int64_t nk, int64_t rb, const double* B,
                                    int64_t sa, int64_t nj, int64_t sb, double* out) {
  if (ra < 1 || ni < 1 || nk < 1 || rb < 1 || sa < 1 || nj < 1 || sb < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(ttm_ttm_core_kernel, dim3(grid_of(ra * sa * ni * nj * rb * sb)), dim3(256),
                     0, (hipStream_t)stream, A, ra, ni, nk, rb, B, sa, nj, sb, out);
  return launch_status();
}

extern "C" int ttx_scale_maxabs_f64(ttx_stream_t stream, int64_t n, double* x) {
  if (n < 0) return TTX_EINVAL;
  if (n == 0) return TTX_OK;
  hipLaunchKernelGGL(scale_maxabs_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, x);
  return launch_status();
}

extern "C" int ttx_truncated_bond_dev_f64(ttx_stream_t stream, const double* S, int64_t len,
                                          const double* nrm, double scale, int64_t max_rank,
                                          int32_t* out) {
  if (len < 1) return TTX_EINVAL;
  hipLaunchKernelGGL(truncated_bond_dev_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, S, len,
                     nrm, scale, max_rank, out);
  return launch_status();
}

extern "C" int ttx_rank_by_threshold_f64(ttx_stream_t stream, const double* S, int64_t len,
                                         double thr, int64_t max_rank, int32_t* out) {
  if (len < 0) return TTX_EINVAL;
  hipLaunchKernelGGL(rank_by_threshold_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, S, len,
                     thr, max_rank, out);
  return launch_status();
}
