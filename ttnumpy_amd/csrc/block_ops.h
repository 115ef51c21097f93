// Block-level (one workgroup) dense FP64 building blocks shared by the
// single-matrix kernels and the batched solver engine.
#pragma once

#include <float.h>

#include "ttx_common.h"

namespace ttx {

constexpr double kEps = DBL_EPSILON;

// Algorithmic flop tally of the block-level routines of one workgroup
// (thread 0 adds; the engine kernels zero it on entry and bank it per slot).
static __shared__ double g_flops;
__device__ __forceinline__ void add_flops(double f) {
  if (threadIdx.x == 0) g_flops += f;
}

// ---------------------------------------------------------------------------
// Batched strided GEMM by one block: C[b] = alpha A[b] B[b] + beta C[b].
// Each wave owns 32x32 output tiles (2x2 MFMA 16x16x4 f64 blocks; blocks
// wholly past M or N are skipped) and reads its operands straight from memory
// with the given strides.  No barrier inside: callers separate dependent
// calls with __syncthreads().
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bgemm(int64_t batch, int64_t M, int64_t N, int64_t K, double alpha,
                                      const double* A, int64_t a_ms, int64_t a_ks, int64_t a_bs,
                                      const double* B, int64_t b_ks, int64_t b_ns, int64_t b_bs,
                                      double beta, double* C, int64_t c_ms, int64_t c_ns,
                                      int64_t c_bs, bool count_flops = true) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int64_t tm = (M + 31) >> 5, tn = (N + 31) >> 5;
  const int64_t per = tm * tn;
  if (count_flops) add_flops(2.0 * (double)batch * (double)M * (double)N * (double)K);
  for (int64_t t = wave; t < batch * per; t += nw) {
    const int64_t bt = t / per, tt = t % per;
    const int64_t m0 = (tt / tn) << 5, n0 = (tt % tn) << 5;
    const double* Ab = A + bt * a_bs;
    const double* Bb = B + bt * b_bs;
    double* Cb = C + bt * c_bs;
    const int64_t ma = m0 + r, mb = m0 + 16 + r, na = n0 + r, nb = n0 + 16 + r;
    const bool va = ma < M, vb = mb < M, wa = na < N, wb = nb < N;
    // 16 x 16 blocks of the tile wholly past M or N are not computed (uniform
    // across the wave; their outputs are never stored)
    const bool m2 = m0 + 16 < M, n2 = n0 + 16 < N;
    double4_t c00 = {0.0, 0.0, 0.0, 0.0}, c01 = c00, c10 = c00, c11 = c00;
    // two k-steps of operands are loaded before their MFMAs (8 loads per
    // lane in flight); the accumulation order is k ascending
    const double* pa0 = Ab + ma * a_ms + kq * a_ks;
    const double* pa1 = Ab + mb * a_ms + kq * a_ks;
    const double* pb0 = Bb + kq * b_ks + na * b_ns;
    const double* pb1 = Bb + kq * b_ks + nb * b_ns;
    const int64_t sa = 4 * a_ks, sb = 4 * b_ks;
    for (int64_t k0 = 0; k0 < K; k0 += 8) {
      double a0[2], a1[2], b0[2], b1[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const bool kin = k0 + 4 * q + kq < K;
        a0[q] = (kin && va) ? pa0[q * sa] : 0.0;
        a1[q] = (kin && vb) ? pa1[q * sa] : 0.0;
        b0[q] = (kin && wa) ? pb0[q * sb] : 0.0;
        b1[q] = (kin && wb) ? pb1[q * sb] : 0.0;
      }
      pa0 += 2 * sa;
      pa1 += 2 * sa;
      pb0 += 2 * sb;
      pb1 += 2 * sb;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (k0 + 4 * q < K) {
          c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[q], b0[q], c00, 0, 0, 0);
          if (n2) c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[q], b1[q], c01, 0, 0, 0);
          if (m2) c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b0[q], c10, 0, 0, 0);
          if (m2 && n2) c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b1[q], c11, 0, 0, 0);
        }
      }
    }
    // D slot (lane, e) of a 16x16 block is row 4e + lane/16, column lane%16
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t row0 = m0 + 4 * e + kq, row1 = row0 + 16;
      if (row0 < M && wa) {
        double* d = Cb + row0 * c_ms + na * c_ns;
        *d = beta == 0.0 ? alpha * c00[e] : alpha * c00[e] + beta * (*d);
      }
      if (row0 < M && wb) {
        double* d = Cb + row0 * c_ms + nb * c_ns;
        *d = beta == 0.0 ? alpha * c01[e] : alpha * c01[e] + beta * (*d);
      }
      if (row1 < M && wa) {
        double* d = Cb + row1 * c_ms + na * c_ns;
        *d = beta == 0.0 ? alpha * c10[e] : alpha * c10[e] + beta * (*d);
      }
      if (row1 < M && wb) {
        double* d = Cb + row1 * c_ms + nb * c_ns;
        *d = beta == 0.0 ? alpha * c11[e] : alpha * c11[e] + beta * (*d);
      }
    }
  }
}

// C[b][m][n] = alpha sum_k A[m][k] B[b][k][n] + beta C[b][m][n] for M <= 16
// with A shared by the batch (the operator-core product of the local matvec:
// M = R n, K = n R'): the batch's columns are numbered as one index
// j = b N + n, so 16 x 32 output tiles run across batch boundaries instead of
// leaving most of a tile per small product empty, and a wave needs half or
// fewer of bgemm's dependent load round trips.  Each output is the same
// MFMA chain (k ascending, four per step) as in bgemm.  No barrier inside.
__device__ __forceinline__ void bgemm_shared_a_m16(int64_t batch, int64_t M, int64_t N, int64_t K,
                                                   double alpha, const double* A, int64_t a_ms,
                                                   int64_t a_ks, const double* B, int64_t b_ks,
                                                   int64_t b_ns, int64_t b_bs, double beta,
                                                   double* C, int64_t c_ms, int64_t c_ns,
                                                   int64_t c_bs, bool count_flops = true) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int64_t cols = batch * N;
  if (count_flops) add_flops(2.0 * (double)batch * (double)M * (double)N * (double)K);
  const bool va = r < M;
  for (int64_t t = wave; t < ((cols + 31) >> 5); t += nw) {
    const int64_t j0 = t << 5;
    const bool n2 = j0 + 16 < cols;  // uniform
    int64_t jb[2];
    bool wv[2];
    const double* pb[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t j = j0 + 16 * h + r;
      wv[h] = j < cols;
      const int64_t bt = wv[h] ? j / N : 0, n = wv[h] ? j - bt * N : 0;
      jb[h] = bt * c_bs + n * c_ns;
      pb[h] = B + bt * b_bs + n * b_ns + kq * b_ks;
    }
    const double* pa = A + r * a_ms + kq * a_ks;
    double4_t c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0;
    const int64_t sa = 4 * a_ks, sb = 4 * b_ks;
    for (int64_t k0 = 0; k0 < K; k0 += 8) {
      double a[2], b0[2], b1[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const bool kin = k0 + 4 * q + kq < K;
        a[q] = (kin && va) ? pa[q * sa] : 0.0;
        b0[q] = (kin && wv[0]) ? pb[0][q * sb] : 0.0;
        b1[q] = (kin && wv[1]) ? pb[1][q * sb] : 0.0;
      }
      pa += 2 * sa;
      pb[0] += 2 * sb;
      pb[1] += 2 * sb;
#pragma unroll
      for (int q = 0; q < 2; ++q)
        if (k0 + 4 * q < K) {
          c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b0[q], c0, 0, 0, 0);
          if (n2) c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b1[q], c1, 0, 0, 0);
        }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t row = 4 * e + kq;
      if (row < M) {
        if (wv[0]) {
          double* d = C + jb[0] + row * c_ms;
          *d = beta == 0.0 ? alpha * c0[e] : alpha * c0[e] + beta * (*d);
        }
        if (wv[1]) {
          double* d = C + jb[1] + row * c_ms;
          *d = beta == 0.0 ? alpha * c1[e] : alpha * c1[e] + beta * (*d);
        }
      }
    }
  }
}

// bgemm in single precision: FP32 operands and FP32 accumulation on
// v_mfma_f32_16x16x4f32.  Its operand layouts are those of the f64
// instruction (A lane l = A[l%16][l/16], B lane l = B[l/16][l%16]) but its
// output is not: lane l element e is D[4 (l/16) + e][l%16] (f64: row
// 4 e + l/16).  Same tiles, strides and load batching as bgemm.  No barrier
// inside.
typedef float floatx4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void bgemm32(int64_t batch, int64_t M, int64_t N, int64_t K, float alpha,
                                        const float* A, int64_t a_ms, int64_t a_ks, int64_t a_bs,
                                        const float* B, int64_t b_ks, int64_t b_ns, int64_t b_bs,
                                        float beta, float* C, int64_t c_ms, int64_t c_ns,
                                        int64_t c_bs) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int64_t tm = (M + 31) >> 5, tn = (N + 31) >> 5;
  const int64_t per = tm * tn;
  add_flops(2.0 * (double)batch * (double)M * (double)N * (double)K);
  for (int64_t t = wave; t < batch * per; t += nw) {
    const int64_t bt = t / per, tt = t % per;
    const int64_t m0 = (tt / tn) << 5, n0 = (tt % tn) << 5;
    const float* Ab = A + bt * a_bs;
    const float* Bb = B + bt * b_bs;
    float* Cb = C + bt * c_bs;
    const int64_t ma = m0 + r, mb = m0 + 16 + r, na = n0 + r, nb = n0 + 16 + r;
    const bool va = ma < M, vb = mb < M, wa = na < N, wb = nb < N;
    const bool m2 = m0 + 16 < M, n2 = n0 + 16 < N;  // as in bgemm
    floatx4_t c00 = {0.0f, 0.0f, 0.0f, 0.0f}, c01 = c00, c10 = c00, c11 = c00;
    const float* pa0 = Ab + ma * a_ms + kq * a_ks;
    const float* pa1 = Ab + mb * a_ms + kq * a_ks;
    const float* pb0 = Bb + kq * b_ks + na * b_ns;
    const float* pb1 = Bb + kq * b_ks + nb * b_ns;
    const int64_t sa = 4 * a_ks, sb = 4 * b_ks;
    for (int64_t k0 = 0; k0 < K; k0 += 8) {
      float a0[2], a1[2], b0[2], b1[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const bool kin = k0 + 4 * q + kq < K;
        a0[q] = (kin && va) ? pa0[q * sa] : 0.0f;
        a1[q] = (kin && vb) ? pa1[q * sa] : 0.0f;
        b0[q] = (kin && wa) ? pb0[q * sb] : 0.0f;
        b1[q] = (kin && wb) ? pb1[q * sb] : 0.0f;
      }
      pa0 += 2 * sa;
      pa1 += 2 * sa;
      pb0 += 2 * sb;
      pb1 += 2 * sb;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (k0 + 4 * q < K) {
          c00 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[q], b0[q], c00, 0, 0, 0);
          if (n2) c01 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[q], b1[q], c01, 0, 0, 0);
          if (m2) c10 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[q], b0[q], c10, 0, 0, 0);
          if (m2 && n2) c11 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[q], b1[q], c11, 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t row0 = m0 + 4 * kq + e, row1 = row0 + 16;
      if (row0 < M && wa) {
        float* d = Cb + row0 * c_ms + na * c_ns;
        *d = beta == 0.0f ? alpha * c00[e] : alpha * c00[e] + beta * (*d);
      }
      if (row0 < M && wb) {
        float* d = Cb + row0 * c_ms + nb * c_ns;
        *d = beta == 0.0f ? alpha * c01[e] : alpha * c01[e] + beta * (*d);
      }
      if (row1 < M && wa) {
        float* d = Cb + row1 * c_ms + na * c_ns;
        *d = beta == 0.0f ? alpha * c10[e] : alpha * c10[e] + beta * (*d);
      }
      if (row1 < M && wb) {
        float* d = Cb + row1 * c_ms + nb * c_ns;
        *d = beta == 0.0f ? alpha * c11[e] : alpha * c11[e] + beta * (*d);
      }
    }
  }
}

// bgemm32 with BF16 operands: the entries of A and B are rounded to BF16
// (round to nearest even, v_cvt_pk_bf16_f32) as they are loaded and the
// products are accumulated in FP32 on v_mfma_f32_16x16x32_bf16.  Lane l
// supplies the 8 operand values of k = 8 (l/16) + j, j < 8, for row (A) or
// column (B) l % 16 of its 16 x 16 block, the same k slots in A and B;
// output lane l element e = C[4 (l/16) + e][l % 16], as for the FP32 MFMA.
// Each wave owns 32 x 32 output tiles.  No barrier inside.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void bgemm_bf16(int64_t batch, int64_t M, int64_t N, int64_t K,
                                           float alpha, const float* A, int64_t a_ms, int64_t a_ks,
                                           int64_t a_bs, const float* B, int64_t b_ks, int64_t b_ns,
                                           int64_t b_bs, float beta, float* C, int64_t c_ms,
                                           int64_t c_ns, int64_t c_bs) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int64_t tm = (M + 31) >> 5, tn = (N + 31) >> 5;
  const int64_t per = tm * tn;
  add_flops(2.0 * (double)batch * (double)M * (double)N * (double)K);
  for (int64_t t = wave; t < batch * per; t += nw) {
    const int64_t bt = t / per, tt = t % per;
    const int64_t m0 = (tt / tn) << 5, n0 = (tt % tn) << 5;
    const float* Ab = A + bt * a_bs;
    const float* Bb = B + bt * b_bs;
    float* Cb = C + bt * c_bs;
    const int64_t ma = m0 + r, mb = m0 + 16 + r, na = n0 + r, nb = n0 + 16 + r;
    const bool va = ma < M, vb = mb < M, wa = na < N, wb = nb < N;
    const bool m2 = m0 + 16 < M, n2 = n0 + 16 < N;
    floatx4_t c00 = {0.0f, 0.0f, 0.0f, 0.0f}, c01 = c00, c10 = c00, c11 = c00;
    for (int64_t k0 = 0; k0 < K; k0 += 32) {
      float fa0[8], fa1[8], fb0[8], fb1[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t kk = k0 + 8 * kq + j;
        const bool kin = kk < K;
        fa0[j] = (kin && va) ? Ab[ma * a_ms + kk * a_ks] : 0.0f;
        fa1[j] = (kin && vb) ? Ab[mb * a_ms + kk * a_ks] : 0.0f;
        fb0[j] = (kin && wa) ? Bb[kk * b_ks + na * b_ns] : 0.0f;
        fb1[j] = (kin && wb) ? Bb[kk * b_ks + nb * b_ns] : 0.0f;
      }
      bf16x8_t a0, a1, b0, b1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        a0[j] = (__bf16)fa0[j];
        a1[j] = (__bf16)fa1[j];
        b0[j] = (__bf16)fb0[j];
        b1[j] = (__bf16)fb1[j];
      }
      c00 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, c00, 0, 0, 0);
      if (n2) c01 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, c01, 0, 0, 0);
      if (m2) c10 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, c10, 0, 0, 0);
      if (m2 && n2) c11 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c11, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t row0 = m0 + 4 * kq + e, row1 = row0 + 16;
      if (row0 < M && wa) {
        float* d = Cb + row0 * c_ms + na * c_ns;
        *d = beta == 0.0f ? alpha * c00[e] : alpha * c00[e] + beta * (*d);
      }
      if (row0 < M && wb) {
        float* d = Cb + row0 * c_ms + nb * c_ns;
        *d = beta == 0.0f ? alpha * c01[e] : alpha * c01[e] + beta * (*d);
      }
      if (row1 < M && wa) {
        float* d = Cb + row1 * c_ms + na * c_ns;
        *d = beta == 0.0f ? alpha * c10[e] : alpha * c10[e] + beta * (*d);
      }
      if (row1 < M && wb) {
        float* d = Cb + row1 * c_ms + nb * c_ns;
        *d = beta == 0.0f ? alpha * c11[e] : alpha * c11[e] + beta * (*d);
      }
    }
  }
}

// y = alpha * M x + beta * y for a row-major M (row stride ld); one wave per
// row so the reads of a row coalesce.  No barrier inside.
__device__ __forceinline__ void bgemv(int64_t rows, int64_t cols, double alpha, const double* M,
                                      int64_t ld, const double* x, double beta, double* y) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  add_flops(2.0 * (double)rows * (double)cols);
  for (int64_t r = wave; r < rows; r += nw) {
    const double* mr = M + r * ld;
    double s = 0.0;
    for (int64_t c = lane; c < cols; c += 64) s += mr[c] * x[c];
    s = wave_sum(s);
    if (lane == 0) y[r] = beta == 0.0 ? alpha * s : alpha * s + beta * y[r];
  }
}

// ---------------------------------------------------------------------------
// Householder QR on a column-major m x n matrix W (leading dim m) held in
// global memory.  On return W holds R in its upper triangle and the reflector
// vectors below the diagonal; tau[j] the reflector scales.  Q (m x k,
// column-major, leading dim m) is formed in Qw.  Block-cooperative.
// ---------------------------------------------------------------------------
// Blocked form (compact WY, LAPACK's geqrf/larft/larfb scheme restated for
// one block): nb-column panels are factored in LDS, T (nb x nb, upper) is
// built from V^T V, and the trailing columns take C -= V (T^T (V^T C)) as two
// bgemm calls per 64-column chunk instead of one pass per column; Q is formed
// backwards the same way.  W keeps R above and the reflectors below the
// diagonal as in the unblocked form.  lds: lds_cap doubles; the panel must fit
// (m nb + 16 * 64 + 272 doubles with nb >= 4), else the unblocked form runs.
constexpr int64_t kQrChunk = 64;
// The blocked form runs for k >= kQrBlockMin with the panel width chosen
// against kQrPanelCap doubles whatever LDS the caller has beyond that, so
// that a given shape takes the same arithmetic in every kernel that calls it.
constexpr int64_t kQrBlockMin = 32;
constexpr int64_t kQrPanelCap = 9600;

__device__ __forceinline__ void qr_wy_block(int64_t m, int64_t j0, int64_t jb, const double* Vp,
                                            int64_t mp, double* T, const double* tau, double* red) {
  // G = V^T V (jb x jb) by waves, then T by thread 0 (dlarft, forward, columnwise)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  double* G = T + 16 * 16;  // 16 x 16
  for (int64_t p = wave; p < jb * jb; p += nwave) {
    const int64_t i = p / jb, c = p - i * jb;
    double d = 0.0;
    if (i < c) {
      for (int64_t r = lane; r < mp; r += 64) d += Vp[i * mp + r] * Vp[c * mp + r];
      d = wave_sum(d);
    }
    if (lane == 0) G[i * 16 + c] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int64_t c = 0; c < jb; ++c) {
      const double tc = tau[j0 + c];
      T[c * 16 + c] = tc;
      for (int64_t i = 0; i < c; ++i) {
        double acc = 0.0;
        for (int64_t l = i; l < c; ++l) acc += T[i * 16 + l] * G[l * 16 + c];
        T[i * 16 + c] = -tc * acc;
      }
      for (int64_t i = c + 1; i < 16; ++i) T[i * 16 + c] = 0.0;
    }
  }
  __syncthreads();
  (void)m;
  (void)red;
}

// C (mp x cols, column stride ldc) <- (I - V op(T) V^T) C, op(T) = T^T when trans
__device__ __forceinline__ void qr_wy_apply(int64_t mp, int64_t jb, const double* Vp, const double* T,
                                            double* C, int64_t ldc, int64_t cols, bool trans,
                                            double* Y) {
  for (int64_t c0 = 0; c0 < cols; c0 += kQrChunk) {
    const int64_t cw = (cols - c0) < kQrChunk ? (cols - c0) : kQrChunk;
    double* Cc = C + c0 * ldc;
    // Y (jb x cw, row-major) = V^T C
    bgemm(1, jb, cw, mp, 1.0, Vp, mp, 1, 0, Cc, 1, ldc, 0, 0.0, Y, cw, 1, 0, false);
    __syncthreads();
    // Y <- op(T) Y: each thread forms its (at most 8) outputs, then all write
    {
      double o[8];
      const int64_t tot = jb * cw;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t e = threadIdx.x + (int64_t)q * blockDim.x;
        double acc = 0.0;
        if (e < tot) {
          const int64_t i = e / cw, c = e - i * cw;
          for (int64_t t = 0; t < jb; ++t) {
            const double tv = trans ? T[t * 16 + i] : T[i * 16 + t];  // T upper
            acc += tv * Y[t * cw + c];
          }
        }
        o[q] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t e = threadIdx.x + (int64_t)q * blockDim.x;
        if (e < tot) Y[e] = o[q];
      }
    }
    __syncthreads();
    // C -= V Y
    bgemm(1, mp, cw, jb, -1.0, Vp, 1, mp, 0, Y, cw, 1, 0, 1.0, Cc, 1, ldc, 0, false);
    __syncthreads();
  }
}

__device__ __forceinline__ void block_householder_qr(int64_t m, int64_t n, double* W, double* tau, double* Qw,
                                     double* red, bool form_q = true, double* lds = nullptr,
                                     int64_t lds_cap = 0) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  const int64_t k = m < n ? m : n;
  __shared__ double s_tau;
  {
    const double dm = (double)m, dn = (double)n, dk = (double)k;
    add_flops(2.0 * dm * dn * dk - (dm + dn) * dk * dk + 2.0 * dk * dk * dk / 3.0);
    if (form_q) add_flops(4.0 * dm * dk * dk - 4.0 * dk * dk * dk / 3.0);
  }
  const int64_t cap = lds_cap < kQrPanelCap ? lds_cap : kQrPanelCap;
  int64_t nbq = 16;
  while (nbq >= 4 && m * nbq + 16 * kQrChunk + 2 * 256 > cap) nbq /= 2;
  if (lds != nullptr && nbq >= 4 && k >= kQrBlockMin) {
    double* Vp = lds;                     // m x nb, column-major (leading dim mp)
    double* T = lds + m * nbq;            // 16 x 16, then G 16 x 16
    double* Y = T + 2 * 256;              // nb x kQrChunk
    for (int64_t j0 = 0; j0 < k; j0 += nbq) {
      const int64_t jb = (k - j0) < nbq ? (k - j0) : nbq;
      const int64_t mp = m - j0;
      for (int64_t e = tid; e < jb * mp; e += blockDim.x) {
        const int64_t c = q32(e, mp), r = e - c * mp;
        Vp[c * mp + r] = W[(j0 + c) * m + j0 + r];
      }
      __syncthreads();
      // unblocked factorization of the panel in LDS
      for (int64_t c = 0; c < jb; ++c) {
        double* col = Vp + c * mp;
        double sq = 0.0;
        for (int64_t i = c + 1 + tid; i < mp; i += blockDim.x) sq += col[i] * col[i];
        sq = block_sum(sq, red);
        const double alpha = col[c];
        __syncthreads();
        if (sq == 0.0) {
          if (tid == 0) { s_tau = 0.0; tau[j0 + c] = 0.0; }
        } else {
          const double beta = -copysign(sqrt(alpha * alpha + sq), alpha);
          const double scal = 1.0 / (alpha - beta);
          for (int64_t i = c + 1 + tid; i < mp; i += blockDim.x) col[i] *= scal;
          if (tid == 0) {
            s_tau = (beta - alpha) / beta;
            tau[j0 + c] = s_tau;
            col[c] = beta;
          }
        }
        __syncthreads();
        const double t = s_tau;
        if (t != 0.0) {
          for (int64_t cc = c + 1 + wave; cc < jb; cc += nwave) {
            double* pc = Vp + cc * mp;
            double d = 0.0;
            for (int64_t i = c + 1 + lane; i < mp; i += 64) d += col[i] * pc[i];
            d = wave_sum(d) + pc[c];
            d *= t;
            if (lane == 0) pc[c] -= d;
            for (int64_t i = c + 1 + lane; i < mp; i += 64) pc[i] -= d * col[i];
          }
        }
        __syncthreads();
      }
      // R and the reflectors back to W; V made explicit (unit diagonal, zeros above)
      for (int64_t e = tid; e < jb * mp; e += blockDim.x) {
        const int64_t c = q32(e, mp), r = e - c * mp;
        W[(j0 + c) * m + j0 + r] = Vp[c * mp + r];
      }
      __syncthreads();
      for (int64_t e = tid; e < jb * mp; e += blockDim.x) {
        const int64_t c = q32(e, mp), r = e - c * mp;
        if (r <= c) Vp[c * mp + r] = r == c ? 1.0 : 0.0;
      }
      __syncthreads();
      if (j0 + jb < n) {
        qr_wy_block(m, j0, jb, Vp, mp, T, tau, red);
        qr_wy_apply(mp, jb, Vp, T, W + (j0 + jb) * m + j0, m, n - j0 - jb, true, Y);
      }
    }
    if (!form_q) return;
    // Q = H_0 ... H_{k-1} [I_k; 0]: block reflectors applied from the last panel back
    for (int64_t e = tid; e < m * k; e += blockDim.x) {
      const int64_t c = q32(e, m), i = e - c * m;
      Qw[e] = (i == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int64_t last = ((k - 1) / nbq) * nbq;
    for (int64_t j0 = last; j0 >= 0; j0 -= nbq) {
      const int64_t jb = (k - j0) < nbq ? (k - j0) : nbq;
      const int64_t mp = m - j0;
      for (int64_t e = tid; e < jb * mp; e += blockDim.x) {
        const int64_t c = q32(e, mp), r = e - c * mp;
        Vp[c * mp + r] = r < c ? 0.0 : (r == c ? 1.0 : W[(j0 + c) * m + j0 + r]);
      }
      __syncthreads();
      qr_wy_block(m, j0, jb, Vp, mp, T, tau, red);
      qr_wy_apply(mp, jb, Vp, T, Qw + j0 * m + j0, m, k - j0, false, Y);
    }
    return;
  }
  for (int64_t j = 0; j < k; ++j) {
    double* col = W + j * m;
    double s = 0.0;
    for (int64_t i = j + 1 + tid; i < m; i += blockDim.x) s += col[i] * col[i];
    s = block_sum(s, red);
    const double alpha = col[j];
    __syncthreads();
    if (s == 0.0) {
      if (tid == 0) { s_tau = 0.0; tau[j] = 0.0; }
    } else {
      const double beta = -copysign(sqrt(alpha * alpha + s), alpha);
      const double scal = 1.0 / (alpha - beta);
      for (int64_t i = j + 1 + tid; i < m; i += blockDim.x) col[i] *= scal;
      if (tid == 0) {
        s_tau = (beta - alpha) / beta;
        tau[j] = s_tau;
        col[j] = beta;
      }
    }
    __syncthreads();
    const double t = s_tau;
    if (t != 0.0) {
      for (int64_t c = j + 1 + wave; c < n; c += nwave) {
        double* cc = W + c * m;
        double d = 0.0;
        for (int64_t i = j + 1 + lane; i < m; i += 64) d += col[i] * cc[i];
        d = wave_sum(d) + cc[j];
        d *= t;
        if (lane == 0) cc[j] -= d;
        for (int64_t i = j + 1 + lane; i < m; i += 64) cc[i] -= d * col[i];
      }
    }
    __syncthreads();
  }
  if (!form_q) return;
  // Q = H_0 ... H_{k-1} [I_k; 0], accumulated backwards
  for (int64_t e = tid; e < m * k; e += blockDim.x) {
    const int64_t c = q32(e, m), i = e - c * m;
    Qw[e] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int64_t j = k - 1; j >= 0; --j) {
    const double t = tau[j];
    const double* v = W + j * m;
    if (t != 0.0) {
      for (int64_t c = j + wave; c < k; c += nwave) {
        double* qc = Qw + c * m;
        double d = 0.0;
        for (int64_t i = j + 1 + lane; i < m; i += 64) d += v[i] * qc[i];
        d = wave_sum(d) + qc[j];
        d *= t;
        if (lane == 0) qc[j] -= d;
        for (int64_t i = j + 1 + lane; i < m; i += 64) qc[i] -= d * v[i];
      }
    }
    __syncthreads();
  }
}

struct SvdArgs {
  int64_t m, n;
  const double* A;
  int64_t a_rs, a_cs;
  double* U;
  int64_t u_rs, u_cs;
  double* S;
  double* Vt;
  int64_t v_rs, v_cs;
  double* work;
  int32_t* info;
  const int* skip = nullptr;  // when non-null and *skip == 0 the kernel returns at once
};



// One rotation of the one-sided Jacobi method on columns i < j of G (and of
// the accumulated V) by the 16 lanes of a DPP row; `sub` is the lane within
// the row.  Element r of a column is handled by lane r % 16, so the sums and
// hence the result do not depend on which row or workgroup size runs it.
// Returns 0 when the pair was not rotated, 2 when it was and either |c|
// exceeded thr sqrt(a b) or the rotation's sine exceeded thr, else 1.  With REGS, columns of up to
// 16 * kJacRegs elements are held in registers between the sums and the
// rotation (the same operations in the same order, so the same bits).
constexpr int kJacRegs = 8;  // column elements per lane held in registers
// the rotated pair (cs x - sn y, sn x + cs y) with the roundings spelled out,
// so that every code path that rotates produces the same bits
__device__ __forceinline__ double jac_rot_lo(double cs, double sn, double x, double y) {
  return __fma_rn(cs, x, -__dmul_rn(sn, y));
}
__device__ __forceinline__ double jac_rot_hi(double cs, double sn, double x, double y) {
  return __fma_rn(sn, x, __dmul_rn(cs, y));
}
// Rotation of a column pair with squared norms a, b and inner product c
// (Rutishauser's formulas, every rounding explicit); false when the pair is
// orthogonal to tol.
__device__ __forceinline__ bool jac_params(double a, double b, double c, double tol, double& cs,
                                           double& sn) {
  if (c == 0.0 || a == 0.0 || b == 0.0) return false;
  if (fabs(c) <= __dmul_rn(__dmul_rn(tol, sqrt(a)), sqrt(b))) return false;
  const double zeta = (b - a) / (2.0 * c);
  const double tn = fabs(zeta) > 1e150
                        ? 0.5 / zeta
                        : (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(__fma_rn(zeta, zeta, 1.0)));
  cs = 1.0 / sqrt(__fma_rn(tn, tn, 1.0));
  sn = __dmul_rn(cs, tn);
  return true;
}

template <bool REGS, bool ACCV>
__device__ __forceinline__ int jacobi_pair(double* gi, double* gj, double* vi, double* vj, int Q,
                                           int sub, double tol, double thr) {
  double a = 0.0, b = 0.0, c = 0.0;
  if (REGS && Q <= 16 * kJacRegs) {
    // loads at clamped (in-bounds) indices, zeros selected past the end
    double x[kJacRegs], y[kJacRegs], u[kJacRegs], w[kJacRegs];
#pragma unroll
    for (int t = 0; t < kJacRegs; ++t) {
      const int r = sub + 16 * t;
      const int rc = r < Q ? r : Q - 1;
      const bool in = r < Q;
      const double gx = gi[rc], gy = gj[rc];
      const double vx = ACCV ? vi[rc] : 0.0, vy = ACCV ? vj[rc] : 0.0;
      x[t] = in ? gx : 0.0;
      y[t] = in ? gy : 0.0;
      u[t] = in ? vx : 0.0;
      w[t] = in ? vy : 0.0;
    }
#pragma unroll
    for (int t = 0; t < kJacRegs; ++t) {
      a = __fma_rn(x[t], x[t], a);
      b = __fma_rn(y[t], y[t], b);
      c = __fma_rn(x[t], y[t], c);
    }
    a = row16_sum(a);
    b = row16_sum(b);
    c = row16_sum(c);
    double cs, sn;
    if (!jac_params(a, b, c, tol, cs, sn)) return 0;
#pragma unroll
    for (int t = 0; t < kJacRegs; ++t) {
      const int r = sub + 16 * t;
      if (r < Q) {
        gi[r] = jac_rot_lo(cs, sn, x[t], y[t]);
        gj[r] = jac_rot_hi(cs, sn, x[t], y[t]);
        if (ACCV) {
          vi[r] = jac_rot_lo(cs, sn, u[t], w[t]);
          vj[r] = jac_rot_hi(cs, sn, u[t], w[t]);
        }
      }
    }
    return (fabs(c) > __dmul_rn(__dmul_rn(thr, sqrt(a)), sqrt(b)) || fabs(sn) > thr) ? 2 : 1;
  }
  // four elements per lane per step, loads first (same order of operations)
  for (int r0 = sub; r0 < Q; r0 += 64) {
    double x[4], y[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int r = r0 + 16 * t;
      const int rc = r < Q ? r : Q - 1;
      const double gx = gi[rc], gy = gj[rc];
      x[t] = r < Q ? gx : 0.0;
      y[t] = r < Q ? gy : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      a = __fma_rn(x[t], x[t], a);
      b = __fma_rn(y[t], y[t], b);
      c = __fma_rn(x[t], y[t], c);
    }
  }
  a = row16_sum(a);
  b = row16_sum(b);
  c = row16_sum(c);
  double cs, sn;
  if (!jac_params(a, b, c, tol, cs, sn)) return 0;
  for (int r0 = sub; r0 < Q; r0 += 64) {
    double x[4], y[4], u[4], w[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int r = r0 + 16 * t;
      const int rc = r < Q ? r : Q - 1;
      x[t] = gi[rc];
      y[t] = gj[rc];
      u[t] = ACCV ? vi[rc] : 0.0;
      w[t] = ACCV ? vj[rc] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int r = r0 + 16 * t;
      if (r < Q) {
        gi[r] = jac_rot_lo(cs, sn, x[t], y[t]);
        gj[r] = jac_rot_hi(cs, sn, x[t], y[t]);
        if (ACCV) {
          vi[r] = jac_rot_lo(cs, sn, u[t], w[t]);
          vj[r] = jac_rot_hi(cs, sn, u[t], w[t]);
        }
      }
    }
  }
  return (fabs(c) > __dmul_rn(__dmul_rn(thr, sqrt(a)), sqrt(b)) || fabs(sn) > thr) ? 2 : 1;
}

// Sweeps of the one-sided Jacobi method on the Q columns of G (Q x Q,
// column-major) with the rotations accumulated into V (with ACCV), until a
// sweep rotates nothing (true) or 80 sweeps (false).  Without ACCV a sweep
// whose rotations all had |c| <= sqrt(tol)/4 sqrt(a b) and a sine below
// sqrt(tol)/4 also ends it: one-sided Jacobi converges quadratically, so the
// next sweep would find every pair orthogonal to about tol/16 and rotate by
// rounding-level angles at most.  The sine test matters for clustered
// singular values: with a ~ b a small |c| can still give an O(1) angle, the
// pair then converges only linearly, and a stop on |c| alone leaves the
// columns orthogonal to ~sqrt(tol) rather than tol.
// Each round of the round-robin schedule pairs every column once; a 16-lane
// DPP row takes one pair.  `flag` points to two block-shared ints.  Ends synced.
template <bool REGS, bool ACCV>
__device__ __forceinline__ bool jacobi_sweeps(double* G, double* V, int Q, int* flag) {
  const int tid = threadIdx.x;
  const int sub = tid & 15;
  const int row = tid >> 4, nrow = blockDim.x >> 4;
  const int QQ = Q + (Q & 1);
  const double tol = (double)Q * kEps;
  if (Q < 2) return true;
  const double thr = ACCV ? 0.0 : 0.25 * sqrt(tol);
  for (int sweep = 0; sweep < 80; ++sweep) {
    if (tid == 0) {
      flag[0] = 0;
      flag[1] = 0;
    }
    __syncthreads();
    for (int round = 0; round < QQ - 1; ++round) {
      for (int pk = row; pk < QQ / 2; pk += nrow) {
        // round-robin: player 0 stays, the others rotate by one per round
        int i = 0, j;
        if (pk != 0) {
          i = pk - 1 + round;
          if (i >= QQ - 1) i -= QQ - 1;
          i += 1;
        }
        j = QQ - 2 - pk + round;
        if (j >= QQ - 1) j -= QQ - 1;
        j += 1;
        if (i >= Q || j >= Q) continue;
        if (i > j) { const int t = i; i = j; j = t; }
        const int rot = jacobi_pair<REGS, ACCV>(G + i * Q, G + j * Q, V + i * Q, V + j * Q, Q, sub,
                                                tol, thr);
        if (rot != 0 && sub == 0) {
          flag[0] = 1;
          if (rot == 2) flag[1] = 1;
        }
      }
      __syncthreads();
    }
    const bool done = flag[0] == 0 || (!ACCV && flag[1] == 0);
    __syncthreads();
    if (done) return true;
  }
  return false;
}

// The columns c of X (rows x cols, X[r * rs + c * cs], rows >= cols) whose
// singular value S[c] is exactly zero are replaced by unit vectors
// orthogonal to every other column, so that an SVD with zero singular
// values still has orthonormal singular vectors, as LAPACK's gesdd/gesvd
// give (the sweeps here leave such columns zero).  The other columns must be
// orthonormal; S is descending, so the zero ones come last.  For each zero
// column, in order, the candidate is the e_j least inside the span of the
// columns before it (the smallest sum_q X[j][q]^2; its distance from the span
// is at least sqrt((rows - c) / rows)), orthogonalized twice and normalized.
// Block-wide (red: blockDim.x / 64 doubles); a single read of S[cols - 1]
// when no singular value is zero.  Ends synced.
__device__ void complete_zero_columns(int64_t rows, int64_t cols, double* X, int64_t rs, int64_t cs,
                                      const double* S, double* red) {
  __shared__ double s_val[16];
  __shared__ int64_t s_idx[16];
  if (cols < 1 || rows < cols || !(S[cols - 1] == 0.0)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  int64_t c0 = cols;
  while (c0 > 0 && S[c0 - 1] == 0.0) --c0;
  for (int64_t c = c0; c < cols; ++c) {
    // the least-covered row j: argmin over rows of sum_{q < c} X[j][q]^2
    double best = INFINITY;
    int64_t bj = rows;
    for (int64_t j = tid; j < rows; j += blockDim.x) {
      double sq = 0.0;
      for (int64_t q = 0; q < c; ++q) {
        const double v = X[j * rs + q * cs];
        sq += v * v;
      }
      if (sq < best) { best = sq; bj = j; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int64_t oj = __shfl_xor(bj, off, 64);
      if (ob < best || (ob == best && oj < bj)) { best = ob; bj = oj; }
    }
    __syncthreads();
    if (lane == 0) { s_val[wave] = best; s_idx[wave] = bj; }
    __syncthreads();
    best = s_val[0];
    bj = s_idx[0];
    for (int w = 1; w < nwave; ++w)
      if (s_val[w] < best || (s_val[w] == best && s_idx[w] < bj)) { best = s_val[w]; bj = s_idx[w]; }
    // v = e_bj, orthogonalized twice against the columns before c, normalized
    for (int64_t r = tid; r < rows; r += blockDim.x) X[r * rs + c * cs] = (r == bj) ? 1.0 : 0.0;
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass)
      for (int64_t q = 0; q < c; ++q) {
        double d = 0.0;
        for (int64_t r = tid; r < rows; r += blockDim.x) d += X[r * rs + q * cs] * X[r * rs + c * cs];
        d = block_sum(d, red);
        for (int64_t r = tid; r < rows; r += blockDim.x) X[r * rs + c * cs] -= d * X[r * rs + q * cs];
        __syncthreads();
      }
    double nn = 0.0;
    for (int64_t r = tid; r < rows; r += blockDim.x) {
      const double v = X[r * rs + c * cs];
      nn += v * v;
    }
    nn = sqrt(block_sum(nn, red));
    for (int64_t r = tid; r < rows; r += blockDim.x) X[r * rs + c * cs] /= nn;
    __syncthreads();
  }
}

// Thin SVD of one matrix by the whole block: QR of the tall side, then
// one-sided Jacobi on the square factor (round-robin pairs, one 16-lane DPP
// row per pair, the square factor in LDS when `lds` holds it), columns sorted
// by descending singular value, and the tall factor's product on MFMA.
// `work` holds 2*P*Q + 2*Q*Q + 3*Q doubles (P >= Q the long and short
// sides); it may be LDS.  REGS: see jacobi_pair.  Without ACCV the rotations
// are not accumulated: V (Vt of a tall, U of a wide input) is left unwritten
// for the caller to form.  Ends synced.
// PRE (with ACCV, on tall factors P > Q with Q >= kSvdPreMinQ; smaller factors
// converge in few sweeps either way and keep the plain order): Jacobi
// preconditioned as in Drmac
// and Veselic's one-sided Jacobi SVD: the columns of X are taken in order of
// decreasing norm (X Pi = Qx R), and the rotations act on the columns of R^T,
// which for graded matrices are far closer to orthogonal than R's (about a
// third of the sweeps).  With R^T V = Y (orthogonal columns, norms sigma):
// X = (Qx V) Sigma (Pi Y Sigma^-1)^T, so U_X = Qx V and V_X = Pi Y Sigma^-1.
constexpr int64_t kSvdPreMinQ = 64;
template <bool REGS = true, bool ACCV = true, bool PRE = false>
__device__ __forceinline__ void block_svd(const SvdArgs& p, double* work, double* red,
                                          double* lds = nullptr, int64_t lds_cap = 0) {
  __shared__ int s_rot[2];
  __shared__ int s_bad;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  const bool wide = p.m < p.n;
  // X = A (tall) or A^T (wide): P x Q with P >= Q
  const int64_t P = wide ? p.n : p.m;
  const int64_t Q = wide ? p.m : p.n;
  const int64_t x_rs = wide ? p.a_cs : p.a_rs;
  const int64_t x_cs = wide ? p.a_rs : p.a_cs;
  // Jacobi on the Q x Q factor and U = Q_x U_r (the QR of the tall side counts itself)
  add_flops(20.0 * (double)Q * (double)Q * (double)Q + 2.0 * (double)P * (double)Q * (double)Q);
  double* Wq = work;              // P x Q   (QR working matrix; later the sorted factor)
  double* tau = Wq + P * Q;       // Q
  double* Qx = tau + Q;           // P x Q   (orthonormal factor)
  double* Gw = Qx + P * Q;        // Q x Q   (Jacobi matrix, when not in LDS)
  double* V = Gw + Q * Q;         // Q x Q   (accumulated rotations)
  double* sig = V + Q * Q;        // Q
  double* ordd = sig + Q;         // Q       (sort permutation, stored as doubles)

  const bool pre = PRE && ACCV && P > Q && Q >= kSvdPreMinQ;
  if (tid == 0) s_bad = 0;
  if (pre) {
    // squared column norms (in ordd), then the source column of each position
    // (in sig): decreasing norm, ties by index; a non-finite norm sorts last
    for (int64_t j = wave; j < Q; j += nwave) {
      double s = 0.0;
      for (int64_t r = lane; r < P; r += 64) {
        const double v = p.A[r * x_rs + j * x_cs];
        s += v * v;
      }
      s = wave_sum(s);
      if (lane == 0) ordd[j] = isfinite(s) ? s : -1.0;
    }
    __syncthreads();
    for (int64_t j = tid; j < Q; j += blockDim.x) {
      const double sj = ordd[j];
      int64_t pos = 0;
      for (int64_t l = 0; l < Q; ++l) {
        const double sl = ordd[l];
        if (sl > sj || (sl == sj && l < j)) ++pos;
      }
      sig[pos] = (double)j;
    }
  }
  __syncthreads();
  for (int64_t e = tid; e < P * Q; e += blockDim.x) {
    const int64_t j = q32(e, P), i = e - j * P;
    const int64_t src = pre ? (int64_t)sig[j] : j;
    const double v = p.A[i * x_rs + src * x_cs];
    if (!isfinite(v)) s_bad = 1;
    Wq[e] = v;
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) p.info[0] = TTX_ENONFINITE;
    return;
  }
  const bool use_qr = P > Q;
  if (use_qr) block_householder_qr(P, Q, Wq, tau, Qx, red, true, lds, lds_cap);
  // the QR is done with the LDS: the Jacobi matrix moves there if it fits
  double* G = (lds != nullptr && Q * Q <= lds_cap) ? lds : Gw;
  if (pre) {
    // G = R^T; the permutation moves to tau (free once Qx is formed)
    for (int64_t e = tid; e < Q * Q; e += blockDim.x) {
      const int64_t c = q32(e, Q), i = e - c * Q;
      G[e] = (c <= i) ? Wq[c + i * P] : 0.0;
    }
    for (int64_t j = tid; j < Q; j += blockDim.x) tau[j] = sig[j];
  } else if (use_qr) {
    for (int64_t e = tid; e < Q * Q; e += blockDim.x) {
      const int64_t c = q32(e, Q), i = e - c * Q;
      G[e] = (i <= c) ? Wq[i + c * P] : 0.0;
    }
  } else {
    for (int64_t e = tid; e < Q * Q; e += blockDim.x) G[e] = Wq[e];
  }
  if (ACCV)
    for (int64_t e = tid; e < Q * Q; e += blockDim.x) V[e] = (e == q32(e, Q) * (Q + 1)) ? 1.0 : 0.0;
  __syncthreads();

  // one-sided Jacobi on the columns of G; a separate call per location of G
  // lets each access compile to an LDS or a global load
  const int converged = G == lds ? jacobi_sweeps<REGS, ACCV>(lds, V, (int)Q, s_rot)
                                 : jacobi_sweeps<REGS, ACCV>(Gw, V, (int)Q, s_rot);
  if (tid == 0) p.info[0] = converged ? TTX_OK : TTX_ENOCONV;

  // singular values = column norms
  for (int64_t j = wave; j < Q; j += nwave) {
    const double* gj = G + j * Q;
    double s = 0.0;
    for (int64_t r = lane; r < Q; r += 64) s += gj[r] * gj[r];
    s = wave_sum(s);
    if (lane == 0) sig[j] = sqrt(s);
  }
  __syncthreads();
  // descending order, stable on ties (selection by rank count)
  for (int64_t j = tid; j < Q; j += blockDim.x) {
    const double sj = sig[j];
    int64_t pos = 0;
    for (int64_t l = 0; l < Q; ++l) {
      const double sl = sig[l];
      if (sl > sj || (sl == sj && l < j)) ++pos;
    }
    ordd[pos] = (double)j;
  }
  __syncthreads();
  // Gs (in Wq) column o = G[:, ord[o]] / sig[ord[o]]  (pre: V[:, ord[o]])
  double* Gs = Wq;
  for (int64_t e = tid; e < Q * Q; e += blockDim.x) {
    const int64_t o = q32(e, Q), r = e - o * Q;
    const int64_t src = (int64_t)ordd[o];
    const double s = sig[src];
    if (pre)
      Gs[e] = V[r + src * Q];
    else
      Gs[e] = s > 0.0 ? G[r + src * Q] / s : 0.0;
  }
  __syncthreads();
  // Ux (P x Q) = (use_qr ? Qx : I) Gs;  tall: U_A = Ux, Vt_A = V^T;  wide: U_A = V, Vt_A = Ux^T
  double* ux = wide ? p.Vt : p.U;
  const int64_t ux_rs = wide ? p.v_cs : p.u_rs, ux_cs = wide ? p.v_rs : p.u_cs;
  if (use_qr) {
    bgemm(1, P, Q, Q, 1.0, Qx, 1, P, 0, Gs, 1, Q, 0, 0.0, ux, ux_rs, ux_cs, 0, false);
  } else {
    for (int64_t e = tid; e < P * Q; e += blockDim.x) {
      const int64_t o = q32(e, P), r = e - o * P;
      ux[r * ux_rs + o * ux_cs] = Gs[r + o * Q];
    }
  }
  if (ACCV)
    for (int64_t e = tid; e < Q * Q; e += blockDim.x) {
      const int64_t o = q32(e, Q), j = e - o * Q;
      const int64_t src = (int64_t)ordd[o];
      double v;
      int64_t r = j;
      if (pre) {
        // row Pi[j] of V_X = Y[j, ord[o]] / sigma
        const double s = sig[src];
        v = s > 0.0 ? G[j + src * Q] / s : 0.0;
        r = (int64_t)tau[j];
      } else {
        v = V[j + src * Q];
      }
      if (!wide)
        p.Vt[o * p.v_rs + r * p.v_cs] = v;
      else
        p.U[r * p.u_rs + o * p.u_cs] = v;
    }
  for (int64_t o = tid; o < Q; o += blockDim.x) p.S[o] = sig[(int64_t)ordd[o]];
  __syncthreads();
  if (ACCV) {
    // zero singular values: complete U's columns and Vt's rows to orthonormal sets
    complete_zero_columns(p.m, Q, p.U, p.u_rs, p.u_cs, p.S, red);
    complete_zero_columns(p.n, Q, p.Vt, p.v_cs, p.v_rs, p.S, red);
  }
}

// ---------------------------------------------------------------------------
// Blocked LU with partial pivoting and the solve of [M | b] by one 512-thread
// block (ttnumpy/solvers/common.py:767-773 solves the same system with
// LAPACK's getrf/getrs).
//
// W is row-major with row stride ld >= n + 1; column n holds b.  Row
// interchanges are logical: perm[i] is the storage row of logical row i, so a
// pivot swaps two ints instead of two rows, and every access goes through
// perm.  Pivots are chosen as getrf does (largest |entry|, first on ties, in
// the current logical order).  Because b rides along as column n, the
// updates leave y = L^{-1} P b there and L is never needed again, so the
// columns left of the current block are not permuted.
//
// Blocking: outer panels of kLuOuter = 8 * nb columns; inside, nb-column
// panels (nb <= 16, staged in LDS column-major) are factored right-looking
// with their updates confined to the outer panel and to its block row; the
// rest of the matrix takes one rank-kLuOuter update per outer panel through
// the LDS-staged MFMA tile routine, which reads and writes the trailing matrix
// once per 128 columns instead of once per 16.
//
// lds: lds_bytes of shared memory (the panel, the GEMM staging, the back
// substitution's diagonal blocks) with perm in its last 4 n bytes; needs
// lds_bytes >= 4 n + 8 max(n, 4224).  NT = blockDim.x, 256 or 512.  *info = TTX_ESINGULAR on a zero pivot
// (the factorization continues, as getrf does).  Ends synced.
// ---------------------------------------------------------------------------
constexpr int kLuTileM = 128;  // trailing-update output tile rows
constexpr int kLuK = 16;       // K slab staged per step

// For logical rows i < Mr and columns j < Nc:
//   W[perm[r0+i]][c0+j] -= sum_{t<K} W[perm[r0+i]][a0+t] * W[perm[b0+t]][c0+j]
// NT threads (256 or 512): output tiles of 128 x (64 NT / 256); each wave owns
// 32 x 64 (2 x 4 MFMA 16x16 blocks) accumulated in registers; A and B are
// staged kLuK deep in LDS, k-major, with the next slab's global loads in
// flight during the MFMAs.  Uses kLuK * (128 + tile columns) doubles of lds.
// Ends synced.
template <int NT>
__device__ __forceinline__ void lu_tile_update_t(int64_t Mr, int64_t Nc, int64_t K, double* W,
                                                 int64_t ld, const int* perm, int64_t r0,
                                                 int64_t a0, int64_t b0, int64_t c0, double* lds) {
  constexpr int TN = 64 * NT / 256;       // tile columns
  constexpr int EA = kLuTileM * kLuK / NT;  // A slab elements per thread (consecutive k)
  constexpr int TPA = kLuK / EA;            // threads per A row
  constexpr int TPB = TN / 4;               // threads per B k-row (4 columns each)
  if (Mr <= 0 || Nc <= 0 || K <= 0) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int wm = (wave & 3) * 32, wn = (wave >> 2) * 64;
  double* As = lds;                     // [kLuK][kLuTileM]
  double* Bs = lds + kLuK * kLuTileM;   // [kLuK][TN]
  const int ar = tid / TPA, ak = (tid % TPA) * EA;
  const int bk = tid / TPB, bc = (tid % TPB) * 4;
  const int64_t tm = (Mr + kLuTileM - 1) / kLuTileM, tn = (Nc + TN - 1) / TN;
  for (int64_t t = 0; t < tm * tn; ++t) {
    const int64_t m0 = (t / tn) * kLuTileM, n0 = (t % tn) * TN;
    // the 8 output rows of this lane: row m0 + wm + 16 mt + 4 e + lq
    double* crow[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = m0 + wm + 16 * mt + 4 * e + lq;
        crow[mt][e] = row < Mr ? W + (int64_t)perm[r0 + row] * ld + c0 : nullptr;
      }
    double4_t acc[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const int64_t col = n0 + wn + 16 * nt + l16;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[mt][nt][e] = (crow[mt][e] && col < Nc) ? crow[mt][e][col] : 0.0;
      }
    const int64_t arow = m0 + ar;
    const double* ap = arow < Mr ? W + (int64_t)perm[r0 + arow] * ld + a0 : nullptr;
    double ra[EA], rb[4];
    auto gload = [&](int64_t k0) {
#pragma unroll
      for (int i = 0; i < EA; ++i) {
        const int64_t kk = k0 + ak + i;
        ra[i] = (ap && kk < K) ? ap[kk] : 0.0;
      }
      const int64_t kb = k0 + bk;
      const double* bp = kb < K ? W + (int64_t)perm[b0 + kb] * ld + c0 + n0 : nullptr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t cc = bc + i;
        rb[i] = (bp && n0 + cc < Nc) ? bp[cc] : 0.0;
      }
    };
    gload(0);
    for (int64_t k0 = 0; k0 < K; k0 += kLuK) {
      __syncthreads();  // the previous slab's MFMAs are done with As/Bs
#pragma unroll
      for (int i = 0; i < EA; ++i) As[(ak + i) * kLuTileM + ar] = ra[i];
#pragma unroll
      for (int i = 0; i < 4; ++i) Bs[bk * TN + bc + i] = rb[i];
      __syncthreads();
      if (k0 + kLuK < K) gload(k0 + kLuK);
#pragma unroll
      for (int ks = 0; ks < kLuK; ks += 4) {
        const int kr = ks + lq;
        double a[2], b[4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) a[mt] = -As[kr * kLuTileM + wm + 16 * mt + l16];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) b[nt] = Bs[kr * TN + wn + 16 * nt + l16];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int nt = 0; nt < 4; ++nt)
            acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const int64_t col = n0 + wn + 16 * nt + l16;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (crow[mt][e] && col < Nc) crow[mt][e][col] = acc[mt][nt][e];
      }
  }
  __syncthreads();
}

// Rows k..k+nb (logical) of columns [cb, ce): x <- L_kk^{-1} x with the unit
// lower L_kk given as L[l * ls + i] = L_kk[i][l] (i > l).  A thread per column,
// so the row reads coalesce.  nb <= 16.  No barrier inside.
__device__ __forceinline__ void lu_trsm_rows(int64_t k, int64_t nb, int64_t cb, int64_t ce, double* W,
                                             int64_t ld, const int* perm, const double* L,
                                             int64_t ls) {
  for (int64_t c = cb + threadIdx.x; c < ce; c += blockDim.x) {
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = i < nb ? W[(int64_t)perm[k + i] * ld + c] : 0.0;
#pragma unroll
    for (int i = 1; i < 16; ++i) {
      if (i < nb) {
        double s = x[i];
#pragma unroll
        for (int l = 0; l < i; ++l) s -= L[l * ls + i] * x[l];
        x[i] = s;
      }
    }
#pragma unroll
    for (int i = 1; i < 16; ++i)
      if (i < nb) W[(int64_t)perm[k + i] * ld + c] = x[i];
  }
}

// clk (may be null): wall-clock ticks added per phase: [0] panels, [1] the
// in-panel triangular solves and updates, [2] block rows, [3] trailing updates,
// [4] back substitution.
__device__ __forceinline__ void lu_clock(int64_t* clk, int i, int64_t& t0) {
  if (clk == nullptr) return;
  const int64_t t = wall_clock64();
  if (threadIdx.x == 0) clk[i] += t - t0;
  t0 = t;
}

template <int NT>
__device__ __forceinline__ void block_lu_solve(int64_t n, int64_t ld, double* W, double* u,
                                               double* lds, int64_t lds_bytes, int32_t* info,
                                               int64_t* clk = nullptr) {
  int64_t t0 = wall_clock64();
  __shared__ double rv[16];
  __shared__ int ri[16];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  int* perm = reinterpret_cast<int*>(reinterpret_cast<char*>(lds) + lds_bytes) - n;
  int64_t nb0 = (lds_bytes - 4 * n) / (8 * n);
  if (nb0 > 16) nb0 = 16;
  const int64_t nbo = 8 * nb0;
  {
    const double dn = (double)n;
    add_flops(2.0 * dn * dn * dn / 3.0 + 2.0 * dn * dn);
  }
  for (int64_t i = tid; i < n; i += blockDim.x) perm[i] = (int)i;
  __syncthreads();
  for (int64_t K0 = 0; K0 < n; K0 += nbo) {
    const int64_t ko = (n - K0) < nbo ? (n - K0) : nbo;
    const int64_t K1 = K0 + ko;
    for (int64_t k = K0; k < K1; k += nb0) {
      const int64_t nb = (K1 - k) < nb0 ? (K1 - k) : nb0;
      const int64_t mp = n - k;
      double* P = lds;
      if ((tid & 15) < nb) {
        // 16 lanes per row (one 128-byte read per row), 8 rows in flight per lane
        const int c = tid & 15;
        const int64_t rs = blockDim.x >> 4;
        for (int64_t rb = tid >> 4; rb < mp; rb += 8 * rs) {
          double v[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int64_t r = rb + q * rs;
            v[q] = r < mp ? W[(int64_t)perm[k + r] * ld + (k + c)] : 0.0;
          }
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int64_t r = rb + q * rs;
            if (r < mp) P[c * mp + r] = v[q];
          }
        }
      }
      __syncthreads();
      for (int64_t j = 0; j < nb; ++j) {
        double* pj = P + j * mp;
        double best = -1.0;
        int bi = (int)j;
        for (int64_t r = j + tid; r < mp; r += blockDim.x) {
          const double v = fabs(pj[r]);
          if (v > best) { best = v; bi = (int)r; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          const double ov = __shfl_xor(best, off, 64);
          const int oi = __shfl_xor(bi, off, 64);
          if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if (lane == 0) { rv[wave] = best; ri[wave] = bi; }
        __syncthreads();
        double b0 = rv[0];
        int pr = ri[0];
        for (int w = 1; w < nwave; ++w)
          if (rv[w] > b0 || (rv[w] == b0 && ri[w] < pr)) { b0 = rv[w]; pr = ri[w]; }
        if (tid == 0) {
          if (b0 == 0.0) info[0] = TTX_ESINGULAR;
          if (pr != j) {
            const int t = perm[k + j];
            perm[k + j] = perm[k + pr];
            perm[k + pr] = t;
          }
        }
        if (pr != j && tid < nb) {
          double* pc = P + tid * mp;
          const double t = pc[j];
          pc[j] = pc[pr];
          pc[pr] = t;
        }
        __syncthreads();
        const double dj = pj[j];
        for (int64_t r = j + 1 + tid; r < mp; r += blockDim.x) {
          double l = pj[r];
          if (dj != 0.0) {
            l /= dj;
            pj[r] = l;
          }
          for (int64_t c = j + 1; c < nb; ++c) P[c * mp + r] -= l * P[c * mp + j];
        }
        __syncthreads();
      }
      if ((tid & 15) < nb) {
        const int c = tid & 15;
        for (int64_t r = tid >> 4; r < mp; r += blockDim.x >> 4)
          W[(int64_t)perm[k + r] * ld + (k + c)] = P[c * mp + r];
      }
      lu_clock(clk, 0, t0);
      // U rows k..k+nb of the panel columns right of this sub-panel, then the
      // rank-nb update of the rest of the outer panel (all rows below)
      lu_trsm_rows(k, nb, k + nb, K1, W, ld, perm, P, mp);
      __syncthreads();
      lu_tile_update_t<NT>(n - (k + nb), K1 - (k + nb), nb, W, ld, perm, k + nb, k, k, k + nb, lds);
      lu_clock(clk, 1, t0);
    }
    // the outer panel is final (its pivots are all known), so its block row
    // U12 = L11^{-1} A12 can be formed: per sub-panel a triangular solve with
    // L_kk (staged in LDS) and a rank-nb update of the block rows below it
    for (int64_t k = K0; k < K1; k += nb0) {
      const int64_t nb = (K1 - k) < nb0 ? (K1 - k) : nb0;
      double* L = lds;  // nb x nb, L[l * 16 + i] = L_kk[i][l]
      for (int64_t e = tid; e < nb * nb; e += blockDim.x) {
        const int64_t i = q32(e, nb), l = e - i * nb;
        L[l * 16 + i] = W[(int64_t)perm[k + i] * ld + k + l];
      }
      __syncthreads();
      lu_trsm_rows(k, nb, K1, n + 1, W, ld, perm, L, 16);
      __syncthreads();
      lu_tile_update_t<NT>(K1 - (k + nb), n + 1 - K1, nb, W, ld, perm, k + nb, k, k, K1, lds);
    }
    lu_clock(clk, 2, t0);
    // the rest of the matrix: one rank-ko update
    lu_tile_update_t<NT>(n - K1, n + 1 - K1, ko, W, ld, perm, K1, K0, K0, K1, lds);
    lu_clock(clk, 3, t0);
  }
  // back substitution U u = y in 64-column blocks from the bottom; the
  // diagonal block is solved by one wave, the rows above are updated one
  // wave per row so each row's reads coalesce
  double* T = lds;  // 64 x 65
  double* xb = lds + 64 * 65;
  for (int64_t i = tid; i < n; i += blockDim.x) u[i] = W[(int64_t)perm[i] * ld + n];
  __syncthreads();
  for (int64_t j1 = n; j1 > 0; j1 -= 64) {
    const int64_t j0 = j1 - 64 > 0 ? j1 - 64 : 0;
    const int64_t nbk = j1 - j0;
    if (lane < nbk) {
      for (int64_t r = wave; r < nbk; r += nwave)
        T[r * 65 + lane] = W[(int64_t)perm[j0 + r] * ld + j0 + lane];
    }
    __syncthreads();
    if (wave == 0) {
      const int64_t row = j0 + lane;
      double yv = row < j1 ? u[row] : 0.0;
      for (int64_t j = j1 - 1; j >= j0; --j) {
        const int jj = (int)(j - j0);
        const double xj = __shfl(yv, jj, 64) / T[jj * 65 + jj];
        if (row == j) yv = xj;
        if (row < j) yv -= xj * T[lane * 65 + jj];
      }
      if (row < j1) { u[row] = yv; xb[lane] = yv; }
    }
    __syncthreads();
    const double xl = lane < nbk ? xb[lane] : 0.0;
    for (int64_t r = wave; r < j0; r += 4 * nwave) {  // 4 rows in flight per wave
      double sv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t rq = r + q * nwave;
        sv[q] = (rq < j0 && lane < nbk) ? W[(int64_t)perm[rq] * ld + j0 + lane] * xl : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double t = wave_sum(sv[q]);
        const int64_t rq = r + q * nwave;
        if (lane == 0 && rq < j0) u[rq] -= t;
      }
    }
    __syncthreads();
  }
  lu_clock(clk, 4, t0);
}

// ---------------------------------------------------------------------------
// Blocked Cholesky solve of a symmetric positive definite [M | b] (row-major,
// row stride ld >= n + 1; only the lower triangle of M and column n = b are
// read, L overwrites the lower triangle) by one block of 256 threads, with
// u = M^{-1} b.
//
// Left-looking over 64-column block columns J = [J0, J0 + jb), with
// L_J< = L[J][0:J0] and y_< = y[0:J0] final from the block columns before:
//  1. D = M_JJ - L_J< L_J<^T (its ten lower 16 x 16 blocks on MFMA) and
//     c = b_J - L_J< y_< into LDS;
//  2. D = L_JJ L_JJ^T in LDS, 16 columns at a time: wave 0 (lane = row of the
//     block) factors the 16-column sub-panel with y_J = L_JJ^{-1} c riding
//     along, then the lower blocks right of it take its rank-16 update on
//     MFMA; L_JJ goes to W and y_J to u;
//  3. per tile of 64 NTB rows R below J: V = (M_RJ - L_R< L_J<^T)^T on MFMA
//     into registers (each wave owns 16 NTB rows of R and all of J), then
//     X L_JJ^T = V^T sixteen columns at a time: the columns' update by the
//     solved ones on MFMA (with -L_JJ from LDS as the A operand; V in the
//     accumulator layout is the B operand as it stands), then one lane per
//     row solves against the 16 x 16 diagonal block by substitution through
//     LDS; X = L_RJ is stored.
// Every entry of L is written once and read back only as an operand:
// L_R< once per block column and L_J< once per tile of R, about
// (n^3 / 384)(1 + 64 / RT) doubles over the factorization, besides reading
// M and writing L once (n^2 / 2 doubles each), where a right-looking
// factorization reads and writes the trailing matrix once per panel.
// Each entry sees the operations of the right-looking blocked form in the
// same order: one chain of MFMA steps (four k at a time, k ascending) over
// every 16-column block left of its own, then the 16-column block's
// substitution (a pivot lane's sqrt and divisions inside a diagonal block,
// division by the stored diagonal below it), so the tile shape (NTB) and the
// schedule change no value.  Then L^T u = y from the bottom in 64-row
// blocks.  lds: kCholLds doubles.  *info = TTX_ENOTSPD on a pivot <= 0 or
// not finite: the routine stops early and W is left partly overwritten.
// Ends synced.
// ---------------------------------------------------------------------------

constexpr int kCholKs = 16;  // k slab of the MFMA updates
constexpr int kCholLds = 64 * 65 + 64 + kCholKs * 64 + kCholKs * 128;

// v of lane src (uniform across the wave) in every lane
__device__ __forceinline__ double lane_value(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

template <int NT, int NTB = 2>
__device__ __forceinline__ void block_chol_solve(int64_t n, int64_t ld, double* W, double* u,
                                                 double* lds, int32_t* info,
                                                 int64_t* clk = nullptr) {
  static_assert(NT == 256, "four waves: the 16-row blocks of D and of the tiles below");
  static_assert(NTB == 1 || NTB == 2, "rows of a tile below the diagonal per wave: 16 NTB");
  constexpr int KS = kCholKs;
  constexpr int RT = 64 * NTB;            // rows of R per tile
  constexpr int EB = RT * KS / NT;        // L_R< slab: consecutive k of one row per thread
  constexpr int TPB = KS / EB;
  static_assert(4 * 16 * NTB * 17 <= KS * (64 + RT), "substitution rows fit the slab space");
  __shared__ int s_bad;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  int64_t t0 = wall_clock64();
  {
    const double dn = (double)n;
    add_flops(dn * dn * dn / 3.0 + 2.0 * dn * dn);
  }
  double* D = lds;                // 64 x 65: M_JJ, then L_JJ (lower)
  double* ys = D + 64 * 65;       // 64: c, then y_J
  double* As = ys + 64;           // [KS][64]: slab of L_J<
  double* Bs = As + KS * 64;      // [KS][RT]: slab of L_R< (step 1: of y_<)
  double* Sw = As + wave * 16 * NTB * 17;  // step 3: this wave's rows, 16 columns (stride 17)
  if (tid == 0) s_bad = 0;
  constexpr int EA = 64 * KS / NT;        // L_J< slab: 64 rows x 4 threads x EA k
  static_assert(KS / EA == 4, "four threads per row of the L_J< slab");
  const int ar = tid >> 2, ak = (tid & 3) * EA;
  const int br = tid / TPB, bk = (tid % TPB) * EB;
  // this wave's lower 16 x 16 blocks of D: w, w + 4, w + 8 of
  // (0,0) (1,0) (1,1) (2,0) (2,1) (2,2) (3,0) (3,1) (3,2) (3,3)
  int qi[3], qj[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int b = wave + 4 * q;
    const int i = b < 1 ? 0 : b < 3 ? 1 : b < 6 ? 2 : 3;
    qi[q] = i;
    qj[q] = b - i * (i + 1) / 2;
  }
  const int nq = wave < 2 ? 3 : 2;  // blocks 0..9
  // this lane's entries of a packed 16 x 16 lower triangle (136 entries):
  // index lane + 64 j is entry (pk_c[j], pk_p[j]); pk_c = -1 past the end
  int pk_c[3], pk_p[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int q = lane + 64 * j;
    int c = 0;
    while ((c + 1) * (c + 2) / 2 <= q && c < 16) ++c;
    pk_c[j] = q < 136 ? c : -1;
    pk_p[j] = q - c * (c + 1) / 2;
  }
  for (int64_t J0 = 0; J0 < n; J0 += 64) {
    const int jb = (int)((n - J0) < 64 ? (n - J0) : 64);
    const double* aprow = ar < jb ? W + (J0 + ar) * ld : nullptr;
    // ---- 1. D = M_JJ - L_J< L_J<^T, c = b_J - L_J< y_<
    {
      double4_t acc[3];
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = 16 * qi[q] + 4 * e + lq, c = 16 * qj[q] + l16;
          acc[q][e] = (q < nq && r < jb && c <= r) ? W[(J0 + r) * ld + J0 + c] : 0.0;
        }
      double cacc = (wave == 0 && lane < jb) ? W[(J0 + lane) * ld + n] : 0.0;
      double ra[EA], ry = 0.0;
      auto gload = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < EA; ++i) {
          const int64_t kk = k0 + ak + i;
          ra[i] = (aprow && kk < J0) ? aprow[kk] : 0.0;
        }
        if (tid < KS) ry = (k0 + tid < J0) ? u[k0 + tid] : 0.0;
      };
      if (J0 > 0) gload(0);
      for (int64_t k0 = 0; k0 < J0; k0 += KS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < EA; ++i) As[(ak + i) * 64 + ar] = ra[i];
        if (tid < KS) Bs[tid] = ry;
        __syncthreads();
        if (k0 + KS < J0) gload(k0 + KS);
#pragma unroll
        for (int ks = 0; ks < KS; ks += 4) {
          const int kr = ks + lq;
#pragma unroll
          for (int q = 0; q < 3; ++q)
            if (q < nq) {  // uniform across the wave
              const double a = -As[kr * 64 + 16 * qi[q] + l16];
              const double b = As[kr * 64 + 16 * qj[q] + l16];
              acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
            }
        }
        if (wave == 0) {
#pragma unroll
          for (int kk = 0; kk < KS; ++kk) cacc -= As[kk * 64 + lane] * Bs[kk];
        }
      }
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (q < nq) {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            D[(16 * qi[q] + 4 * e + lq) * 65 + 16 * qj[q] + l16] = acc[q][e];
        }
      if (wave == 0) ys[lane] = cacc;
    }
    __syncthreads();
    lu_clock(clk, 1, t0);
    // ---- 2. D = L_JJ L_JJ^T, y_J = L_JJ^{-1} c
    for (int c0 = 0; c0 < jb; c0 += 16) {
      const int nb = jb - c0 < 16 ? jb - c0 : 16;
      if (wave == 0) {
        // lane r >= c0 holds D[r][c0 .. c0 + 16) (on or below the diagonal);
        // rows of the diagonal block divide by the pivot's sqrt, the rows
        // below it by the stored diagonal entry
        const bool act = lane >= c0 && lane < jb;
        const bool below = lane >= c0 + 16;
        double row[16];
#pragma unroll
        for (int c = 0; c < 16; ++c)
          row[c] = (act && c < nb && c0 + c <= lane) ? D[lane * 65 + c0 + c] : 0.0;
        double yv = act ? ys[lane] : 0.0;
        bool bad = false;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          if (j < nb) {  // uniform
            const int pj = c0 + j;
            const double djj = lane_value(row[j], pj);
            bad = bad || !(djj > 0.0) || !(djj < INFINITY);
            const double d = sqrt(djj);
            const double ljj = djj / d;  // the pivot lane's own row[j] / d
            const double lj = lane >= pj ? row[j] / (below ? ljj : d) : 0.0;
            if (lane >= pj) row[j] = lj;
            const double yj = lane_value(yv, pj) / d;
            if (lane == pj) yv = yj;
#pragma unroll
            for (int c = j + 1; c < 16; ++c)
              if (c < nb) {
                const double lcj = lane_value(lj, c0 + c);
                if (lane > pj && c0 + c <= lane) row[c] -= lj * lcj;
              }
            if (lane > pj) yv -= lj * yj;
          }
        }
        if (act) {
#pragma unroll
          for (int c = 0; c < 16; ++c)
            if (c < nb && c0 + c <= lane) D[lane * 65 + c0 + c] = row[c];
          ys[lane] = yv;
        }
        if (lane == 0 && bad) s_bad = 1;
      }
      __syncthreads();
      if (s_bad) {
        if (tid == 0) info[0] = TTX_ENOTSPD;
        return;
      }
      // rank-16 update of D's lower blocks right of this sub-panel
      const int sb = c0 >> 4;
      if (c0 + 16 < jb) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (q < nq && qj[q] > sb) {  // uniform
            double4_t acc;
#pragma unroll
            for (int e = 0; e < 4; ++e)
              acc[e] = D[(16 * qi[q] + 4 * e + lq) * 65 + 16 * qj[q] + l16];
#pragma unroll
            for (int ks = 0; ks < 16; ks += 4) {
              const double a = -D[(16 * qi[q] + l16) * 65 + c0 + ks + lq];
              const double b = D[(16 * qj[q] + l16) * 65 + c0 + ks + lq];
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
              D[(16 * qi[q] + 4 * e + lq) * 65 + 16 * qj[q] + l16] = acc[e];
          }
        __syncthreads();
      }
    }
    // L_JJ to W, y_J to u
    for (int p = tid; p < 64 * 64; p += NT) {
      const int r = p >> 6, c = p & 63;
      if (r < jb && c <= r) W[(J0 + r) * ld + J0 + c] = D[r * 65 + c];
    }
    if (tid < jb) u[J0 + tid] = ys[tid];
    lu_clock(clk, 0, t0);
    // ---- 3. the rows below: X L_JJ^T = M_RJ - L_R< L_J<^T
    for (int64_t R0 = J0 + 64; R0 < n; R0 += RT) {
      const int64_t wr0 = R0 + wave * 16 * NTB;
      // lane l's rows of the tile (null past n); its columns are J0 + 16 mt + 4 e + l / 16
      double* rp[NTB];
#pragma unroll
      for (int nt = 0; nt < NTB; ++nt) {
        const int64_t row = wr0 + 16 * nt + l16;
        rp[nt] = row < n ? W + row * ld + J0 + lq : nullptr;
      }
      double4_t v[4][NTB];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            v[mt][nt][e] = (rp[nt] && 16 * mt + 4 * e + lq < jb) ? rp[nt][16 * mt + 4 * e] : 0.0;
      const double* bprow = (R0 + br) < n ? W + (R0 + br) * ld : nullptr;
      double ra[EA], rb[EB];
      auto gload = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < EA; ++i) {
          const int64_t kk = k0 + ak + i;
          ra[i] = (aprow && kk < J0) ? aprow[kk] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < EB; ++i) {
          const int64_t kk = k0 + bk + i;
          rb[i] = (bprow && kk < J0) ? bprow[kk] : 0.0;
        }
      };
      if (J0 > 0) gload(0);
      for (int64_t k0 = 0; k0 < J0; k0 += KS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < EA; ++i) As[(ak + i) * 64 + ar] = ra[i];
#pragma unroll
        for (int i = 0; i < EB; ++i) Bs[(bk + i) * RT + br] = rb[i];
        __syncthreads();
        if (k0 + KS < J0) gload(k0 + KS);
#pragma unroll
        for (int ks = 0; ks < KS; ks += 4) {
          const int kr = ks + lq;
          double a[4], b[NTB];
#pragma unroll
          for (int mt = 0; mt < 4; ++mt) a[mt] = -As[kr * 64 + 16 * mt + l16];
#pragma unroll
          for (int nt = 0; nt < NTB; ++nt) b[nt] = Bs[kr * RT + wave * 16 * NTB + 16 * nt + l16];
#pragma unroll
          for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NTB; ++nt)
              v[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt], b[nt], v[mt][nt], 0, 0, 0);
        }
      }
      __syncthreads();  // the slab space becomes this step's substitution rows
      // 16 columns at a time: lane l's element e of block (mt, nt) is
      // V[16 mt + 4 e + l / 16][16 nt + l % 16], the B operand of k-step e
      // of block column mt
      const double* dl = D + l16 * 65 + lq;
#pragma unroll
      for (int m2 = 0; m2 < 4; ++m2) {
#pragma unroll
        for (int mt = 0; mt < m2; ++mt)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double a = -dl[16 * m2 * 65 + 16 * mt + 4 * e];
#pragma unroll
            for (int nt = 0; nt < NTB; ++nt)
              v[m2][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, v[mt][nt][e], v[m2][nt], 0, 0, 0);
          }
        if (16 * m2 < jb) {  // uniform
#pragma unroll
          for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e) Sw[(16 * nt + l16) * 17 + 4 * e + lq] = v[m2][nt][e];
          // the lower triangle of the diagonal block L_bb, packed (entry
          // (c, p) at c (c + 1) / 2 + p) three entries a lane: read from LDS
          // once, then each one broadcast by v_readlane where the solve uses it
          double lt[3];
#pragma unroll
          for (int j = 0; j < 3; ++j)
            lt[j] = pk_c[j] >= 0 ? D[(16 * m2 + pk_c[j]) * 65 + 16 * m2 + pk_p[j]] : 0.0;
          __builtin_amdgcn_wave_barrier();
          {
            // row lane % (16 NTB) of the wave: x L_bb^T = w against the
            // diagonal block.  Every lane runs it (lanes past 16 NTB repeat a
            // row and store the same bits): the entries are read from other
            // lanes, so no lane may skip the code that forms them.
            double* sr = Sw + (lane & (16 * NTB - 1)) * 17;
            double x[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
              double t = sr[c];
#pragma unroll
              for (int p = 0; p < c; ++p) {
                const int q = c * (c + 1) / 2 + p;
                t -= x[p] * lane_value(lt[q >> 6], q & 63);
              }
              const int qd = c * (c + 1) / 2 + c;
              x[c] = 16 * m2 + c < jb ? t / lane_value(lt[qd >> 6], qd & 63) : 0.0;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) sr[c] = x[c];
          }
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e) v[m2][nt][e] = Sw[(16 * nt + l16) * 17 + 4 * e + lq];
        }
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (rp[nt] && 16 * mt + 4 * e + lq < jb) rp[nt][16 * mt + 4 * e] = v[mt][nt][e];
    }
    __syncthreads();  // D is read above until here
    lu_clock(clk, 3, t0);
  }
  // L^T u = y, 64-row blocks from the bottom (y is in u)
  double* T = lds;  // 64 x 65: T[r][c] = L[j0 + r][j0 + c], r >= c
  double* ub = lds + 64 * 65;
  for (int64_t j1 = n; j1 > 0; j1 -= 64) {
    const int64_t j0 = j1 - 64 > 0 ? j1 - 64 : 0;
    const int64_t nbk = j1 - j0;
    {
      // the block's 16 rows of this wave: loads in flight together, then stored
      double tv[64 * 64 / NT];
#pragma unroll
      for (int j = 0; j < 64 * 64 / NT; ++j) {
        const int r = wave + (NT / 64) * j;
        tv[j] = (lane < nbk && r < nbk && lane <= r) ? W[(j0 + r) * ld + j0 + lane] : 0.0;
      }
      if (lane < nbk) {
#pragma unroll
        for (int j = 0; j < 64 * 64 / NT; ++j) {
          const int r = wave + (NT / 64) * j;
          if (r < nbk) T[r * 65 + lane] = tv[j];
        }
      }
    }
    __syncthreads();
    if (wave == 0) {
      const int64_t row = j0 + lane;
      double yv = row < j1 ? u[row] : 0.0;
      for (int jj = (int)nbk - 1; jj >= 0; --jj) {
        const double xj = lane_value(yv, jj) / T[jj * 65 + jj];
        if (lane == jj) yv = xj;
        if (lane < jj) yv -= xj * T[jj * 65 + lane];  // L[j][row] = L^T[row][j]
      }
      if (row < j1) { u[row] = yv; ub[lane] = yv; }
    }
    __syncthreads();
    // y_i -= sum_{c in block} L[c][i] u_c for the rows above (coalesced over i)
    // (32 loads of a thread in flight; the sum runs over c in order)
    for (int64_t i = tid; i < j0; i += NT) {
      double acc = 0.0;
      for (int64_t c = 0; c < nbk; c += 32) {
        double v[32];
#pragma unroll
        for (int q = 0; q < 32; ++q) v[q] = (c + q < nbk) ? W[(j0 + c + q) * ld + i] : 0.0;
#pragma unroll
        for (int q = 0; q < 32; ++q)
          if (c + q < nbk) acc += v[q] * ub[c + q];
      }
      u[i] -= acc;
    }
    __syncthreads();
  }
  lu_clock(clk, 4, t0);
}

// ---------------------------------------------------------------------------
// FP32 Cholesky factorization (no right-hand side) for the mixed-precision
// local solve: the same blocking as block_chol_solve, in single precision on
// v_mfma_f32_16x16x4f32 (A lane l = A[l%16][l/16], B lane l = B[l/16][l%16],
// output lane l element e = D[4 (l/16) + e][l%16], established on the box by
// an in-bounds probe).  W: row-major, row stride ld, lower triangle read and
// overwritten by L.  *info = TTX_ENOTSPD on a pivot <= 0 or not finite.
// ---------------------------------------------------------------------------
typedef float float4_t __attribute__((ext_vector_type(4)));

// Left-looking, like block_chol_solve: for each 64-column block column J
//   1. D = M_JJ - L_J< L_J<^T on MFMA (ten lower 16 x 16 blocks over four waves)
//   2. D = L_JJ L_JJ^T by wave 0 (lane = row), 16-column sub-panels, the lower
//      blocks right of each sub-panel taking its rank-16 update on MFMA
//   3. each tile of 64 NTB rows R below: V^T = (M_RJ - L_R< L_J<^T)^T on MFMA
//      into registers, then X L_JJ^T = V solved in registers (off-diagonal
//      16-column blocks on MFMA, the diagonal ones by one lane per row).
// Every L entry is written once and the columns left of J are read once per
// tile.  The FP32 accumulator (lane l, element e) holds row 4 (l/16) + e of
// its 16 x 16 block, so in step 3 lane l holds V[row l%16][16 mt + 4 (l/16) + e]:
// as the B operand of k-step e the lanes of quarter q supply column
// 16 mt + 4 q + e, and the A operand (L_JJ) is read with the same k order.
// lds: kChol32Lds floats.  Ends synced.
constexpr int kChol32Ks = 32;  // k slab of the MFMA updates
constexpr int kChol32Lds = 64 * 65 + kChol32Ks * 64 + kChol32Ks * 128;

template <int NT, int NTB = 2>
__device__ __forceinline__ void block_chol32_factor(int64_t n, int64_t ld, float* W, float* lds,
                                                    int32_t* info, int64_t* clk = nullptr) {
  static_assert(NT == 256, "four waves: the 16-row blocks of D and of the tiles below");
  static_assert(NTB == 1 || NTB == 2, "rows of a tile below the diagonal per wave: 16 NTB");
  constexpr int KS = kChol32Ks;
  constexpr int RT = 64 * NTB;            // rows of R per tile
  constexpr int EA = 64 * KS / NT;        // L_J< slab: consecutive k of one row per thread
  constexpr int TPA = KS / EA;
  constexpr int EB = RT * KS / NT;        // L_R< slab
  constexpr int TPB = KS / EB;
  static_assert(TPA == 4, "four threads per row of the L_J< slab");
  static_assert(4 * 16 * NTB * 17 <= KS * (64 + RT), "substitution rows fit the slab space");
  static_assert(64 * 65 + KS * (64 + RT) <= kChol32Lds, "LDS plan");
  __shared__ int s_bad32;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  {
    const double dn = (double)n;
    add_flops(dn * dn * dn / 3.0);
  }
  float* D = lds;                // 64 x 65: M_JJ, then L_JJ (lower)
  float* As = D + 64 * 65;       // [KS][64]: slab of L_J<
  float* Bs = As + KS * 64;      // [KS][RT]: slab of L_R<
  float* Sw = As + wave * 16 * NTB * 17;  // step 3: this wave's rows, 16 columns (stride 17)
  if (tid == 0) s_bad32 = 0;
  const int ar = tid / TPA, ak = (tid % TPA) * EA;
  const int br = tid / TPB, bk = (tid % TPB) * EB;
  // this wave's lower 16 x 16 blocks of D: w, w + 4, w + 8 of
  // (0,0) (1,0) (1,1) (2,0) (2,1) (2,2) (3,0) (3,1) (3,2) (3,3)
  int qi[3], qj[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int b = wave + 4 * q;
    const int i = b < 1 ? 0 : b < 3 ? 1 : b < 6 ? 2 : 3;
    qi[q] = i;
    qj[q] = b - i * (i + 1) / 2;
  }
  const int nq = wave < 2 ? 3 : 2;  // blocks 0..9
  // this lane's entries of a packed 16 x 16 lower triangle (136 entries):
  // index lane + 64 j is entry (pk_c[j], pk_p[j]); pk_c = -1 past the end
  int pk_c[3], pk_p[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int q = lane + 64 * j;
    int c = 0;
    while ((c + 1) * (c + 2) / 2 <= q && c < 16) ++c;
    pk_c[j] = q < 136 ? c : -1;
    pk_p[j] = q - c * (c + 1) / 2;
  }
  int64_t t0 = wall_clock64();
  for (int64_t J0 = 0; J0 < n; J0 += 64) {
    const int jb = (int)((n - J0) < 64 ? (n - J0) : 64);
    const float* aprow = ar < jb ? W + (J0 + ar) * ld : nullptr;
    float ra[EA];
    auto aload = [&](int64_t k0) {
#pragma unroll
      for (int i = 0; i < EA; ++i) {
        const int64_t kk = k0 + ak + i;
        ra[i] = (aprow && kk < J0) ? aprow[kk] : 0.0f;
      }
    };
    // ---- 1. D = M_JJ - L_J< L_J<^T
    {
      float4_t acc[3];
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = 16 * qi[q] + 4 * lq + e, c = 16 * qj[q] + l16;
          acc[q][e] = (q < nq && r < jb && c <= r) ? W[(J0 + r) * ld + J0 + c] : 0.0f;
        }
      if (J0 > 0) aload(0);
      for (int64_t k0 = 0; k0 < J0; k0 += KS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < EA; ++i) As[(ak + i) * 64 + ar] = ra[i];
        __syncthreads();
        if (k0 + KS < J0) aload(k0 + KS);
#pragma unroll
        for (int ks = 0; ks < KS; ks += 4) {
          const int kr = ks + lq;
#pragma unroll
          for (int q = 0; q < 3; ++q)
            if (q < nq) {  // uniform across the wave
              const float a = -As[kr * 64 + 16 * qi[q] + l16];
              const float b = As[kr * 64 + 16 * qj[q] + l16];
              acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
            }
        }
      }
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (q < nq) {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            D[(16 * qi[q] + 4 * lq + e) * 65 + 16 * qj[q] + l16] = acc[q][e];
        }
    }
    __syncthreads();
    // ---- 2. D = L_JJ L_JJ^T
    for (int c0 = 0; c0 < jb; c0 += 16) {
      const int nb = jb - c0 < 16 ? jb - c0 : 16;
      if (wave == 0) {
        // lane r >= c0 holds D[r][c0 .. c0 + 16) (on or below the diagonal)
        const bool act = lane >= c0 && lane < jb;
        float row[16];
#pragma unroll
        for (int c = 0; c < 16; ++c)
          row[c] = (act && c < nb && c0 + c <= lane) ? D[lane * 65 + c0 + c] : 0.0f;
        bool bad = false;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          if (j < nb) {  // uniform
            const int pj = c0 + j;
            const float djj = __shfl(row[j], pj, 64);
            bad = bad || !(djj > 0.0f) || !(djj < INFINITY);
            const float d = sqrtf(djj);
            const float lj = lane >= pj ? row[j] / d : 0.0f;
            if (lane >= pj) row[j] = lj;
#pragma unroll
            for (int c = j + 1; c < 16; ++c)
              if (c < nb) {
                const float lcj = __shfl(lj, c0 + c, 64);
                if (lane > pj && c0 + c <= lane) row[c] -= lj * lcj;
              }
          }
        }
        if (act) {
#pragma unroll
          for (int c = 0; c < 16; ++c)
            if (c < nb && c0 + c <= lane) D[lane * 65 + c0 + c] = row[c];
        }
        if (lane == 0 && bad) s_bad32 = 1;
      }
      __syncthreads();
      if (s_bad32) {
        if (tid == 0) info[0] = TTX_ENOTSPD;
        return;
      }
      // rank-16 update of D's lower blocks right of this sub-panel
      const int sb = c0 >> 4;
      if (c0 + 16 < jb) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (q < nq && qj[q] > sb) {  // uniform
            float4_t acc;
#pragma unroll
            for (int e = 0; e < 4; ++e)
              acc[e] = D[(16 * qi[q] + 4 * lq + e) * 65 + 16 * qj[q] + l16];
#pragma unroll
            for (int ks = 0; ks < 16; ks += 4) {
              const float a = -D[(16 * qi[q] + l16) * 65 + c0 + ks + lq];
              const float b = D[(16 * qj[q] + l16) * 65 + c0 + ks + lq];
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
              D[(16 * qi[q] + 4 * lq + e) * 65 + 16 * qj[q] + l16] = acc[e];
          }
        __syncthreads();
      }
    }
    // L_JJ to W
    for (int p = tid; p < 64 * 64; p += NT) {
      const int r = p >> 6, c = p & 63;
      if (r < jb && c <= r) W[(J0 + r) * ld + J0 + c] = D[r * 65 + c];
    }
    lu_clock(clk, 0, t0);
    // ---- 3. the rows below: X L_JJ^T = M_RJ - L_R< L_J<^T
    for (int64_t R0 = J0 + 64; R0 < n; R0 += RT) {
      const int64_t wr0 = R0 + wave * 16 * NTB;
      // lane l's rows of the tile (null past n); its columns are J0 + 16 mt + 4 (l / 16) + e
      float* rp[NTB];
#pragma unroll
      for (int nt = 0; nt < NTB; ++nt) {
        const int64_t row = wr0 + 16 * nt + l16;
        rp[nt] = row < n ? W + row * ld + J0 + 4 * lq : nullptr;
      }
      float4_t v[4][NTB];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            v[mt][nt][e] = (rp[nt] && 16 * mt + 4 * lq + e < jb) ? rp[nt][16 * mt + e] : 0.0f;
      const float* bprow = (R0 + br) < n ? W + (R0 + br) * ld : nullptr;
      float rb[EB];
      auto bload = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < EB; ++i) {
          const int64_t kk = k0 + bk + i;
          rb[i] = (bprow && kk < J0) ? bprow[kk] : 0.0f;
        }
      };
      if (J0 > 0) {
        aload(0);
        bload(0);
      }
      for (int64_t k0 = 0; k0 < J0; k0 += KS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < EA; ++i) As[(ak + i) * 64 + ar] = ra[i];
#pragma unroll
        for (int i = 0; i < EB; ++i) Bs[(bk + i) * RT + br] = rb[i];
        __syncthreads();
        if (k0 + KS < J0) {
          aload(k0 + KS);
          bload(k0 + KS);
        }
#pragma unroll
        for (int ks = 0; ks < KS; ks += 4) {
          const int kr = ks + lq;
          float a[4], b[NTB];
#pragma unroll
          for (int mt = 0; mt < 4; ++mt) a[mt] = -As[kr * 64 + 16 * mt + l16];
#pragma unroll
          for (int nt = 0; nt < NTB; ++nt) b[nt] = Bs[kr * RT + wave * 16 * NTB + 16 * nt + l16];
#pragma unroll
          for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NTB; ++nt)
              v[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mt], b[nt], v[mt][nt], 0, 0, 0);
        }
      }
      __syncthreads();  // the slab space becomes this step's substitution rows
      // k-step e of block column mt: lane l supplies column 16 mt + 4 (l/16) + e
      const float* dl = D + l16 * 65 + 4 * lq;
#pragma unroll
      for (int m2 = 0; m2 < 4; ++m2) {
#pragma unroll
        for (int mt = 0; mt < m2; ++mt)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float a = -dl[16 * m2 * 65 + 16 * mt + e];
#pragma unroll
            for (int nt = 0; nt < NTB; ++nt)
              v[m2][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, v[mt][nt][e], v[m2][nt], 0, 0, 0);
          }
        if (16 * m2 < jb) {  // uniform
#pragma unroll
          for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e) Sw[(16 * nt + l16) * 17 + 4 * lq + e] = v[m2][nt][e];
          // the lower triangle of L_bb, packed (entry (c, p) at c (c + 1) / 2 + p)
          // three entries a lane, the diagonal ones inverted: read from LDS
          // once, then each one broadcast by v_readlane where the solve uses it
          float lt[3];
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const float v = pk_c[j] >= 0 ? D[(16 * m2 + pk_c[j]) * 65 + 16 * m2 + pk_p[j]] : 1.0f;
            lt[j] = pk_c[j] == pk_p[j] ? __builtin_amdgcn_rcpf(v) : v;
          }
          __builtin_amdgcn_wave_barrier();
          {
            // row lane % (16 NTB) of the wave: x L_bb^T = w against the
            // diagonal block.  Every lane runs it (lanes past 16 NTB repeat a
            // row and store the same bits): the entries are read from other
            // lanes, so no lane may skip the code that forms them.
            float* sr = Sw + (lane & (16 * NTB - 1)) * 17;
            float x[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
              float t = sr[c];
#pragma unroll
              for (int p = 0; p < c; ++p) {
                const int q = c * (c + 1) / 2 + p;
                t -= x[p] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lt[q >> 6]), q & 63));
              }
              const int qd = c * (c + 1) / 2 + c;
              const float rd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lt[qd >> 6]), qd & 63));
              x[c] = 16 * m2 + c < jb ? t * rd : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) sr[c] = x[c];
          }
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e) v[m2][nt][e] = Sw[(16 * nt + l16) * 17 + 4 * lq + e];
        }
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTB; ++nt)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (rp[nt] && 16 * mt + 4 * lq + e < jb) rp[nt][16 * mt + e] = v[mt][nt][e];
    }
    __syncthreads();  // D is read above until here
    lu_clock(clk, 1, t0);
  }
}

// Inverses of the 64 x 64 diagonal blocks of the FP32 Cholesky factor L
// (row-major, row stride ld, lower triangle) into Dinv: block b at
// Dinv + 4096 b, row-major 64 x 64, lower; a partial last block is padded
// with the identity.  Two waves take a block (two blocks at a time with four
// waves): column c of the inverse is formed by forward substitution on e_c
// by two lanes of one wave, lane c % 32 holding rows 0..31 and lane
// c % 32 + 32 rows 32..63 in 32 registers each; the pivot of each step is
// passed between them by one shuffle, and every lane runs the same unrolled
// steps (entries left of the diagonal stay zero).  lds: 2 * (64 * 65 + 64)
// floats.  Ends synced.
template <int NT>
__device__ __forceinline__ void chol32_invert_diag(int64_t n, int64_t ld, const float* W, float* Dinv,
                                                   float* lds) {
  static_assert(NT == 256, "two blocks of two waves each");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = NT >> 6;
  const int per = nwave / 2;            // blocks at a time
  const int slot = wave >> 1;           // which of them this wave works on
  const int half = lane >> 5;           // rows 32 * half .. 32 * half + 31
  const int col = (lane & 31) + 32 * (wave & 1);
  const int64_t nblk = (n + 63) / 64;
  float* T = lds + slot * (64 * 65 + 64);  // T[i * 65 + k] = L_bb[i][k]
  float* rd = T + 64 * 65;                 // 1 / L_bb[k][k]
  for (int64_t base = 0; base < nblk; base += per) {
    const int64_t b = base + slot;
    const int64_t j0 = 64 * b;
    const int64_t nbk = b < nblk ? ((n - j0) < 64 ? (n - j0) : 64) : 0;
    if (b < nblk) {
      // 16 loads of this thread in flight, then their LDS stores (twice)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int r = (wave & 1) + 2 * (16 * h + j);
          v[j] = (r < nbk && lane < nbk) ? (lane <= r ? W[(j0 + r) * ld + j0 + lane] : 0.0f)
                                         : (r == lane ? 1.0f : 0.0f);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int r = (wave & 1) + 2 * (16 * h + j);
          T[r * 65 + lane] = v[j];
          if (lane == r) rd[r] = 1.0f / v[j];
        }
      }
    }
    __syncthreads();
    if (b < nblk) {
      const float* Th = T + 32 * half * 65;  // this lane's rows
      float x[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) x[i] = (32 * half + i == col) ? 1.0f : 0.0f;
#pragma unroll
      for (int k = 0; k < 64; ++k) {
        const int owner = k < 32 ? 0 : 1;
        const float piv = x[k & 31] * rd[k];
        const float xk = __shfl(piv, (lane & 31) + 32 * owner, 64);
        if (half == owner) x[k & 31] = xk;
#pragma unroll
        for (int i = 0; i < 32; ++i)
          if (32 * half + i > k) x[i] -= Th[i * 65 + k] * xk;
      }
      float* D = Dinv + b * 4096;
#pragma unroll
      for (int i = 0; i < 32; ++i) D[(32 * half + i) * 64 + col] = x[i];
    }
    __syncthreads();
  }
}

// x <- (L L^T)^{-1} x in FP32 with the factor from block_chol32_factor and
// the diagonal-block inverses from chol32_invert_diag; x in global memory
// (n floats, n <= 4096), the solution so far in LDS.  The row stride ld is a
// multiple of 4, so that 16-byte loads of 4 columns are aligned.  Both sweeps
// step over 64-row blocks and read their rows of L as one pass with 32 floats
// of every lane in flight (16 lanes per 64-column row segment, 4 rows per
// wave per load, 128 rows per pass):
//   forward (right-looking)  y_b = Dinv_b y_b, then y_{>b} -= L[>b, b] y_b
//   back    (left-looking)   z_b = Dinv_b^T (y_b - L[>b, b]^T z_>b)
// The next block's inverse is loaded while the current one is applied.
// lds: 64 x 65 + 4 x 64 + n floats.  Ends synced.
template <int NT>
__device__ __forceinline__ void chol32_apply_inverse(int64_t n, int64_t ld, const float* W,
                                                     const float* Dinv, float* x, float* lds) {
  static_assert(NT == 256, "four waves: 16 rows of a block each, 4 threads per row");
  constexpr int P = 8;  // rows per lane per pass
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4, m = lane & 15;  // row of the wave's four, column quad
  float* D = lds;               // 64 x 65: the block's inverse
  float* part = D + 64 * 65;    // 4 x 64 partial sums of the waves
  float* y = part + 4 * 64;     // n: the solution so far
  const int q = tid & 3, rq = tid >> 2;  // 4 threads per row of the block product
  const int64_t nblk = (n + 63) / 64;
  for (int64_t i = tid; i < n; i += NT) y[i] = x[i];
  float dv[16];
  auto dload = [&](int64_t b) {
#pragma unroll
    for (int j = 0; j < 16; ++j) dv[j] = Dinv[b * 4096 + (wave + 4 * j) * 64 + lane];
  };
  auto dstore = [&]() {
#pragma unroll
    for (int j = 0; j < 16; ++j) D[(wave + 4 * j) * 65 + lane] = dv[j];
  };
  // y_b <- D y_b (trans: D^T y_b), four threads per row; ends synced
  auto dprod = [&](int64_t j0, int64_t nbk, bool trans) {
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int c = 16 * q + t;
      const float dval = trans ? D[c * 65 + rq] : D[rq * 65 + c];
      acc += dval * (c < nbk ? y[j0 + c] : 0.0f);
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    __syncthreads();  // every read of y_b before it is overwritten
    if (q == 0 && rq < nbk) y[j0 + rq] = acc;
    __syncthreads();
  };
  // this lane's 4 columns of rows i0 + 4 p + g (zero past n)
  auto rows_load = [&](int64_t i0, int64_t j0, float4* l) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int64_t i = i0 + 4 * p + g;
      l[p] = i < n ? *reinterpret_cast<const float4*>(W + i * ld + j0 + 4 * m)
                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  };
  dload(0);
  __syncthreads();
  // forward: L y = x
  for (int64_t b = 0; b < nblk; ++b) {
    const int64_t j0 = 64 * b;
    const int64_t nbk = (n - j0) < 64 ? (n - j0) : 64;
    const int64_t j1 = j0 + nbk;
    dstore();
    __syncthreads();
    if (b + 1 < nblk) dload(b + 1);
    dprod(j0, nbk, false);
    if (j1 < n) {  // uniform; the block is full (nbk = 64)
      const float y0 = y[j0 + 4 * m], y1 = y[j0 + 4 * m + 1], y2 = y[j0 + 4 * m + 2],
                  y3 = y[j0 + 4 * m + 3];
      for (int64_t i0 = j1 + 4 * P * wave; i0 < n; i0 += 4 * P * (NT / 64)) {
        float4 l[P];
        rows_load(i0, j0, l);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          float sm = (l[p].x * y0 + l[p].y * y1) + (l[p].z * y2 + l[p].w * y3);
          sm += __shfl_xor(sm, 1, 64);
          sm += __shfl_xor(sm, 2, 64);
          sm += __shfl_xor(sm, 4, 64);
          sm += __shfl_xor(sm, 8, 64);
          const int64_t i = i0 + 4 * p + g;
          if (m == 0 && i < n) y[i] -= sm;
        }
      }
      __syncthreads();
    }
  }
  // back: L^T z = y, blocks from the last
  dload(nblk - 1);
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t j0 = 64 * b;
    const int64_t nbk = (n - j0) < 64 ? (n - j0) : 64;
    const int64_t j1 = j0 + nbk;
    dstore();
    // part[w][c] = sum over the wave's rows i >= j1 of L[i][j0 + c] z_i
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (int64_t i0 = j1 + 4 * P * wave; i0 < n; i0 += 4 * P * (NT / 64)) {
      float4 l[P];
      rows_load(i0, j0, l);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int64_t i = i0 + 4 * p + g;
        const float zi = i < n ? y[i] : 0.0f;
        a0 += l[p].x * zi;
        a1 += l[p].y * zi;
        a2 += l[p].z * zi;
        a3 += l[p].w * zi;
      }
    }
    a0 += __shfl_xor(a0, 16, 64);
    a1 += __shfl_xor(a1, 16, 64);
    a2 += __shfl_xor(a2, 16, 64);
    a3 += __shfl_xor(a3, 16, 64);
    a0 += __shfl_xor(a0, 32, 64);
    a1 += __shfl_xor(a1, 32, 64);
    a2 += __shfl_xor(a2, 32, 64);
    a3 += __shfl_xor(a3, 32, 64);
    if (g == 0) {
      part[wave * 64 + 4 * m] = a0;
      part[wave * 64 + 4 * m + 1] = a1;
      part[wave * 64 + 4 * m + 2] = a2;
      part[wave * 64 + 4 * m + 3] = a3;
    }
    __syncthreads();
    if (b > 0) dload(b - 1);
    if (tid < 64 && tid < nbk)
      y[j0 + tid] -= (part[tid] + part[64 + tid]) + (part[128 + tid] + part[192 + tid]);
    __syncthreads();
    dprod(j0, nbk, true);
  }
  for (int64_t i = tid; i < n; i += NT) x[i] = y[i];
  __syncthreads();
}

}  // namespace ttx
