// Dense building blocks: strided batched FP64 GEMM on MFMA, strided N-d copy,
// fills, axpby and deterministic norms.
//
// The GEMM is the one contraction engine of the hot path: every tensordot of
// the environment recursion, the one-site local apply, the residual core and
// the core absorbs is issued as a strided batched GEMM, so no permuted copy of
// a core or environment is ever materialised.

#include "ttx_common.h"

namespace ttx {

struct GemmArgs {
  int64_t M, N, K;
  double alpha, beta;
  const double* A;
  int64_t a_ms, a_ks, a_bs;
  const double* B;
  int64_t b_ks, b_ns, b_bs;
  double* C;
  int64_t c_ms, c_ns, c_bs;
  const double* skip;   // when set and *skip != 0, the launch does nothing
};

// BM x BN output tile per 256-thread block; the 4 waves split it 2 x 2, each
// owning a (BM/2) x (BN/2) quadrant of (BM/32) x (BN/32) MFMA 16x16 blocks.
// K is staged 16 at a time through LDS, k-major so that the MFMA operand
// reads are consecutive across lanes.  64 x 64 is the default; 32 x 32 (one
// MFMA block per wave, four times the waves and a quarter of the MFMA chain
// each) serves the contractions whose 64 x 64 grid would leave most compute
// units idle.  Both issue the same MFMA sequence per output element, so they
// give identical results.
// Operand layouts of v_mfma_f64_16x16x4f64 (wave64):
//   A (16x4): lane l holds A[l % 16][l / 16]
//   B (4x16): lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, element i holds D[4 * i + l / 16][l % 16]
constexpr int GBM = 64, GBN = 64, GBK = 16;
constexpr int GSM = 32, GSN = 32;

template <int BM, int BN>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmArgs p) {
  if (p.skip != nullptr && *p.skip != 0.0) return;  // uniform across the grid
  constexpr int FM = BM / 32, FN = BN / 32;     // MFMA blocks per wave along m and n
  __shared__ double As[GBK][BM + 1];
  __shared__ double Bs[GBK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t b = blockIdx.z;
  const int64_t m0 = (int64_t)blockIdx.y * BM;
  const int64_t n0 = (int64_t)blockIdx.x * BN;
  const double* A = p.A + b * p.a_bs;
  const double* B = p.B + b * p.b_bs;
  double* C = p.C + b * p.c_bs;
  // pick the fastest-varying tile index so that global reads coalesce
  const bool a_m_fast = p.a_ms <= p.a_ks;
  const bool b_n_fast = p.b_ns <= p.b_ks;
  const int wm = (wave >> 1) * (BM / 2);
  const int wn = (wave & 1) * (BN / 2);
  double4_t acc[FM][FN];
#pragma unroll
  for (int x = 0; x < FM; ++x)
#pragma unroll
    for (int y = 0; y < FN; ++y) acc[x][y] = double4_t{0, 0, 0, 0};
  // each thread stages kPerA / kPerB elements of each slab; the next slab's
  // loads are issued before the current slab's MFMAs
  constexpr int kPerA = BM * GBK / 256, kPerB = BN * GBK / 256;
  int am[kPerA], ak[kPerA], bn[kPerB], bk[kPerB];
#pragma unroll
  for (int q = 0; q < kPerA; ++q) {
    const int e = q * 256 + tid;
    if (a_m_fast) { am[q] = e % BM; ak[q] = e / BM; } else { ak[q] = e % GBK; am[q] = e / GBK; }
  }
#pragma unroll
  for (int q = 0; q < kPerB; ++q) {
    const int e = q * 256 + tid;
    if (b_n_fast) { bn[q] = e % BN; bk[q] = e / BN; } else { bk[q] = e % GBK; bn[q] = e / GBK; }
  }
  double ra[kPerA], rb[kPerB];
  auto load_slab = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < kPerA; ++q) {
      const int64_t gm = m0 + am[q], gka = k0 + ak[q];
      ra[q] = (gm < p.M && gka < p.K) ? A[gm * p.a_ms + gka * p.a_ks] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kPerB; ++q) {
      const int64_t gn = n0 + bn[q], gkb = k0 + bk[q];
      rb[q] = (gn < p.N && gkb < p.K) ? B[gkb * p.b_ks + gn * p.b_ns] : 0.0;
    }
  };
  load_slab(0);
  for (int64_t k0 = 0; k0 < p.K; k0 += GBK) {
#pragma unroll
    for (int q = 0; q < kPerA; ++q) As[ak[q]][am[q]] = ra[q];
#pragma unroll
    for (int q = 0; q < kPerB; ++q) Bs[bk[q]][bn[q]] = rb[q];
    __syncthreads();
    if (k0 + GBK < p.K) load_slab(k0 + GBK);
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      const int kr = kk + (lane >> 4);
      const int c = lane & 15;
      double av[FM], bv[FN];
#pragma unroll
      for (int x = 0; x < FM; ++x) av[x] = As[kr][wm + 16 * x + c];
#pragma unroll
      for (int y = 0; y < FN; ++y) bv[y] = Bs[kr][wn + 16 * y + c];
#pragma unroll
      for (int x = 0; x < FM; ++x)
#pragma unroll
        for (int y = 0; y < FN; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }

  const int col = lane & 15;
  const int rq = lane >> 4;
#pragma unroll
  for (int x = 0; x < FM; ++x)
#pragma unroll
    for (int y = 0; y < FN; ++y) {
      const int64_t gn = n0 + wn + 16 * y + col;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t gm = m0 + wm + 16 * x + 4 * i + rq;
        if (gm < p.M && gn < p.N) {
          double* dst = C + gm * p.c_ms + gn * p.c_ns;
          const double v = p.alpha * acc[x][y][i];
          *dst = (p.beta == 0.0) ? v : v + p.beta * (*dst);
        }
      }
    }
}

struct CopyArgs {
  int ndim;
  int64_t dims[6];
  int64_t ss[6];
  int64_t ds[6];
  int64_t total;
  double alpha;
  const double* src;
  double* dst;
};

__global__ __launch_bounds__(256) void copy_nd_kernel(CopyArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.total; i += stride) {
    int64_t rem = i, so = 0, dof = 0;
    for (int d = a.ndim - 1; d >= 0; --d) {
      const int64_t c = rem % a.dims[d];
      rem /= a.dims[d];
      so += c * a.ss[d];
      dof += c * a.ds[d];
    }
    a.dst[dof] = (a.alpha == 0.0) ? 0.0 : a.alpha * a.src[so];
  }
}

__global__ __launch_bounds__(256) void fill_kernel(double* dst, int64_t n, double v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = v;
}

__global__ __launch_bounds__(256) void axpby_kernel(int64_t n, double a, const double* x, double b,
                                                    const double* y, double* z) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    z[i] = a * x[i] + b * y[i];
}

// Deterministic sum of squares in one 1024-thread block (fixed reduction order).
__global__ __launch_bounds__(1024) void sumsq_single_kernel(int64_t n, const double* a,
                                                           const double* b, double* out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = b ? a[i] - b[i] : a[i];
    s += v * v;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = sqrt(s);
}

static int norm_impl(hipStream_t st, int64_t n, const double* a, const double* b, double* out) {
  if (n < 0) return TTX_EINVAL;
  hipLaunchKernelGGL(sumsq_single_kernel, dim3(1), dim3(1024), 0, st, n, a, b, out);
  return launch_status();
}

static inline unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (unsigned)g;
}

}  // namespace ttx

using namespace ttx;

extern "C" int ttx_version(void) { return 100; }

// FP32 contraction of FP64 arrays: operands rounded to FP32 when staged,
// products accumulated in FP32 on v_mfma_f32_16x16x4f32 (output slot (lane l,
// element e) = row 4 (l/16) + e, column l % 16), the result widened to FP64.
// Same tiling and strides as gemm_f64_kernel.
typedef float float4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void gemm_f32acc_kernel(GemmArgs p, const double* a_scale,
                                                          const double* b_scale) {
  __shared__ float As[GBK][GBM + 1];
  __shared__ float Bs[GBK][GBN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t b = blockIdx.z;
  const int64_t m0 = (int64_t)blockIdx.y * GBM;
  const int64_t n0 = (int64_t)blockIdx.x * GBN;
  const double* A = p.A + b * p.a_bs;
  const double* B = p.B + b * p.b_bs;
  double* C = p.C + b * p.c_bs;
  const bool a_m_fast = p.a_ms <= p.a_ks;
  const bool b_n_fast = p.b_ns <= p.b_ks;
  const int wm = (wave >> 1) * 32;
  const int wn = (wave & 1) * 32;
  // operands are divided by their scales before the rounding to FP32 (so that
  // large trains stay inside FP32's range) and the product multiplied back
  const double sa = (a_scale && *a_scale > 0.0 && isfinite(*a_scale)) ? *a_scale : 1.0;
  const double sb = (b_scale && *b_scale > 0.0 && isfinite(*b_scale)) ? *b_scale : 1.0;
  const double ia = 1.0 / sa, ib = 1.0 / sb;
  float4_t acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
  for (int64_t k0 = 0; k0 < p.K; k0 += GBK) {
#pragma unroll
    for (int e0 = 0; e0 < GBM * GBK; e0 += 256) {
      const int e = e0 + tid;
      int mm, kk;
      if (a_m_fast) { mm = e % GBM; kk = e / GBM; } else { kk = e % GBK; mm = e / GBK; }
      const int64_t gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < p.M && gk < p.K) ? (float)(A[gm * p.a_ms + gk * p.a_ks] * ia) : 0.0f;
    }
#pragma unroll
    for (int e0 = 0; e0 < GBN * GBK; e0 += 256) {
      const int e = e0 + tid;
      int nn, kk;
      if (b_n_fast) { nn = e % GBN; kk = e / GBN; } else { kk = e % GBK; nn = e / GBK; }
      const int64_t gn = n0 + nn, gk = k0 + kk;
      Bs[kk][nn] = (gn < p.N && gk < p.K) ? (float)(B[gk * p.b_ks + gn * p.b_ns] * ib) : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      const int kr = kk + (lane >> 4);
      const int c = lane & 15;
      const float a0 = As[kr][wm + c];
      const float a1 = As[kr][wm + 16 + c];
      const float b0 = Bs[kr][wn + c];
      const float b1 = Bs[kr][wn + 16 + c];
      acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc00, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc01, 0, 0, 0);
      acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc10, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc11, 0, 0, 0);
    }
    __syncthreads();
  }
  const int col = lane & 15;
  const int rq = lane >> 4;
  float4_t* accs[4] = {&acc00, &acc01, &acc10, &acc11};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t gn = n0 + wn + 16 * (q & 1) + col;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t gm = m0 + wm + 16 * (q >> 1) + 4 * rq + i;
      if (gm < p.M && gn < p.N) {
        double* dst = C + gm * p.c_ms + gn * p.c_ns;
        const double v = p.alpha * sa * sb * (double)(*accs[q])[i];
        *dst = (p.beta == 0.0) ? v : v + p.beta * (*dst);
      }
    }
  }
}

extern "C" int ttx_gemm_f32acc(ttx_stream_t stream, int64_t batch, int64_t M, int64_t N, int64_t K,
                               double alpha, const double* A, int64_t a_ms, int64_t a_ks,
                               int64_t a_bs, const double* B, int64_t b_ks, int64_t b_ns,
                               int64_t b_bs, double beta, double* C, int64_t c_ms, int64_t c_ns,
                               int64_t c_bs, const double* a_scale, const double* b_scale) {
  if (batch < 0 || M < 0 || N < 0 || K < 0) return TTX_EINVAL;
  if (batch == 0 || M == 0 || N == 0) return TTX_OK;
  if (batch > 65535) return TTX_EINVAL;
  // the row tiles go on the grid's y dimension (at most 65535): taller
  // products run as several launches over row blocks
  constexpr int64_t kRowsPerLaunch = (int64_t)65535 * GBM;
  if (M > kRowsPerLaunch) {
    for (int64_t r0 = 0; r0 < M; r0 += kRowsPerLaunch) {
      const int64_t mc = (M - r0) < kRowsPerLaunch ? (M - r0) : kRowsPerLaunch;
      const int rc = ttx_gemm_f32acc(stream, batch, mc, N, K, alpha, A + r0 * a_ms, a_ms, a_ks, a_bs,
                                     B, b_ks, b_ns, b_bs, beta, C + r0 * c_ms, c_ms, c_ns, c_bs,
                                     a_scale, b_scale);
      if (rc != TTX_OK) return rc;
    }
    return TTX_OK;
  }
  if (K == 0) {
    int64_t dims[3] = {batch, M, N};
    int64_t ds[3] = {c_bs, c_ms, c_ns};
    return ttx_copy_f64(stream, 3, dims, C, ds, C, ds, beta);
  }
  GemmArgs p{M, N, K, alpha, beta, A, a_ms, a_ks, a_bs, B, b_ks, b_ns, b_bs, C, c_ms, c_ns, c_bs};
  dim3 grid((unsigned)((N + GBN - 1) / GBN), (unsigned)((M + GBM - 1) / GBM), (unsigned)batch);
  hipLaunchKernelGGL(gemm_f32acc_kernel, grid, dim3(256), 0, (hipStream_t)stream, p, a_scale,
                     b_scale);
  return launch_status();
}

// below this many 64 x 64 tiles the 32 x 32 tiling runs; two tiles per compute
// unit unless an autotuner has set another limit (process-wide)
static int64_t g_small_tile_below = 512;

extern "C" int ttx_gemm_set_small_tile_limit(int64_t tiles) {
  if (tiles < 0) return TTX_EINVAL;
  g_small_tile_below = tiles;
  return TTX_OK;
}

extern "C" int64_t ttx_gemm_small_tile_limit(void) { return g_small_tile_below; }

extern "C" int ttx_gemm_f64(ttx_stream_t stream, int64_t batch, int64_t M, int64_t N, int64_t K,
                            double alpha, const double* A, int64_t a_ms, int64_t a_ks,
                            int64_t a_bs, const double* B, int64_t b_ks, int64_t b_ns,
                            int64_t b_bs, double beta, double* C, int64_t c_ms, int64_t c_ns,
                            int64_t c_bs) {
  return gemm_f64_launch(stream, batch, M, N, K, alpha, A, a_ms, a_ks, a_bs, B, b_ks, b_ns, b_bs,
                         beta, C, c_ms, c_ns, c_bs, nullptr);
}

int ttx::gemm_f64_launch(ttx_stream_t stream, int64_t batch, int64_t M, int64_t N, int64_t K,
                    double alpha, const double* A, int64_t a_ms, int64_t a_ks, int64_t a_bs,
                    const double* B, int64_t b_ks, int64_t b_ns, int64_t b_bs, double beta,
                    double* C, int64_t c_ms, int64_t c_ns, int64_t c_bs, const double* skip) {
  if (batch < 0 || M < 0 || N < 0 || K < 0) return TTX_EINVAL;
  if (batch == 0 || M == 0 || N == 0) return TTX_OK;
  if (batch > 65535) return TTX_EINVAL;
  // the row tiles go on the grid's y dimension (at most 65535): taller
  // products run as several launches over row blocks (either tiling gives
  // the same values, so the split does not change them)
  constexpr int64_t kRowsPerLaunch = (int64_t)65535 * GSM;
  if (M > kRowsPerLaunch) {
    for (int64_t r0 = 0; r0 < M; r0 += kRowsPerLaunch) {
      const int64_t mc = (M - r0) < kRowsPerLaunch ? (M - r0) : kRowsPerLaunch;
      const int rc = gemm_f64_launch(stream, batch, mc, N, K, alpha, A + r0 * a_ms, a_ms, a_ks, a_bs,
                                     B, b_ks, b_ns, b_bs, beta, C + r0 * c_ms, c_ms, c_ns, c_bs, skip);
      if (rc != TTX_OK) return rc;
    }
    return TTX_OK;
  }
  if (K == 0) {
    // empty contraction: C = beta * C (beta == 0 writes zeros without reading C)
    int64_t dims[3] = {batch, M, N};
    int64_t ds[3] = {c_bs, c_ms, c_ns};
    return ttx_copy_f64(stream, 3, dims, C, ds, C, ds, beta);
  }
  GemmArgs p{M, N, K, alpha, beta, A, a_ms, a_ks, a_bs, B, b_ks, b_ns, b_bs, C, c_ms, c_ns, c_bs,
              skip};
  const int64_t tiles = ((N + GBN - 1) / GBN) * ((M + GBM - 1) / GBM) * batch;
  if (tiles < g_small_tile_below) {
    dim3 grid((unsigned)((N + GSN - 1) / GSN), (unsigned)((M + GSM - 1) / GSM), (unsigned)batch);
    hipLaunchKernelGGL((gemm_f64_kernel<GSM, GSN>), grid, dim3(256), 0, (hipStream_t)stream, p);
  } else {
    dim3 grid((unsigned)((N + GBN - 1) / GBN), (unsigned)((M + GBM - 1) / GBM), (unsigned)batch);
    hipLaunchKernelGGL((gemm_f64_kernel<GBM, GBN>), grid, dim3(256), 0, (hipStream_t)stream, p);
  }
  return launch_status();
}

extern "C" int ttx_copy_f64(ttx_stream_t stream, int ndim, const int64_t* dims, const double* src,
                            const int64_t* src_strides, double* dst, const int64_t* dst_strides,
                            double alpha) {
  if (ndim < 1 || ndim > 6) return TTX_EINVAL;
  CopyArgs a;
  a.ndim = ndim;
  a.total = 1;
  for (int d = 0; d < 6; ++d) {
    a.dims[d] = d < ndim ? dims[d] : 1;
    a.ss[d] = d < ndim ? src_strides[d] : 0;
    a.ds[d] = d < ndim ? dst_strides[d] : 0;
    if (a.dims[d] < 0) return TTX_EINVAL;
    a.total *= a.dims[d];
  }
  if (a.total == 0) return TTX_OK;
  a.alpha = alpha;
  a.src = src;
  a.dst = dst;
  hipLaunchKernelGGL(copy_nd_kernel, dim3(grid_for(a.total)), dim3(256), 0, (hipStream_t)stream, a);
  return launch_status();
}

extern "C" int ttx_fill_f64(ttx_stream_t stream, double* dst, int64_t n, double value) {
  if (n < 0) return TTX_EINVAL;
  if (n == 0) return TTX_OK;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, dst, n, value);
  return launch_status();
}

extern "C" int ttx_axpby_f64(ttx_stream_t stream, int64_t n, double a, const double* x, double b,
                             const double* y, double* z) {
  if (n < 0) return TTX_EINVAL;
  if (n == 0) return TTX_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, a, x, b, y, z);
  return launch_status();
}

extern "C" int ttx_nrm2_f64(ttx_stream_t stream, int64_t n, const double* x, double* out) {
  return norm_impl((hipStream_t)stream, n, x, nullptr, out);
}

extern "C" int ttx_diff_nrm2_f64(ttx_stream_t stream, int64_t n, const double* a, const double* b,
                                 double* out) {
  return norm_impl((hipStream_t)stream, n, a, b, out);
}
