// Shared helpers of the ttx device library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ttx.h"

namespace ttx {

constexpr int kWave = 64;

// e / m in 32-bit arithmetic for 0 <= e < 2^31, 0 < m < 2^31: 64-bit integer
// division expands to a long instruction sequence on the GPU, and the flat
// index loops over per-request arrays never leave 32-bit range.
__device__ __forceinline__ int64_t q32(int64_t e, int64_t m) {
  return (int64_t)((uint32_t)e / (uint32_t)m);
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// ttx_gemm_f64 with an optional device flag: when `skip` is set and *skip is
// nonzero at run time, the launch does nothing (Krylov steps queued after the
// stop test fired)
int gemm_f64_launch(ttx_stream_t stream, int64_t batch, int64_t M, int64_t N, int64_t K,
                    double alpha, const double* A, int64_t a_ms, int64_t a_ks, int64_t a_bs,
                    const double* B, int64_t b_ks, int64_t b_ns, int64_t b_bs, double beta,
                    double* C, int64_t c_ms, int64_t c_ns, int64_t c_bs, const double* skip);

inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? TTX_OK : TTX_EHIP;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// One DPP lane move of a double within rows of 16 lanes (CTRL is a DPP16
// control word: quad_perm, row_mirror, row_half_mirror).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int mlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const int mhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(mhi, mlo);
}

// Sum over the 16 lanes of a DPP row (lanes 16q..16q+15); every lane of the
// row gets the same bits: each step adds a lane's value to its partner's,
// and a + b == b + a exactly.
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]: partner lane ^ 1
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]: partner lane ^ 2
  v += dpp_move<0x141>(v);  // row_half_mirror: lane i <- 7 - i within 8
  v += dpp_move<0x140>(v);  // row_mirror: lane i <- 15 - i within 16
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// block_sum of v[0..m) at once (m <= M): each value is reduced by the same
// lane tree and wave order as block_sum, behind one pair of barriers.  `red`
// must hold M * blockDim.x / 64 doubles of shared memory.
template <int M>
__device__ __forceinline__ void block_sum_many(double* v, int m, double* red) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwave = (blockDim.x + kWave - 1) / kWave;
#pragma unroll
  for (int j = 0; j < M; ++j)
    if (j < m) v[j] = wave_sum(v[j]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < M; ++j)
      if (j < m) red[wave * M + j] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < M; ++j)
    if (j < m) {
      double t = 0.0;
      for (int w = 0; w < nwave; ++w) t += red[w * M + j];
      v[j] = t;
    }
  __syncthreads();  // red may be reused at once
}

// Sum over the whole block; every thread gets the result.  `red` must hold
// blockDim.x / 64 doubles of shared memory.  Fixed order, so deterministic.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwave = (blockDim.x + kWave - 1) / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nwave; ++w) t += red[w];
  return t;
}

// One entry of the explicit one-site matrix (ttnumpy/solvers/one_site.py:49-75):
//   sum_a psi[t_l, a, f_l] * sum_b A[a, i, j, b] * phi[t_r, b, f_r]
// psi_t = psi + t_l*Ra*fl, A_ij = A + (i*nc + j)*Rb (a stride nr*nc*Rb),
// phi_t = phi + t_r*Rb*fr.  Every caller sums in this order, so the direct
// launches and the engine build bit-identical matrices.
__device__ __forceinline__ double local_matrix_entry(const double* psi_t, int64_t fl, int64_t f_l,
                                                    const double* A_ij, int64_t a_stride, int64_t Ra,
                                                    int64_t Rb, const double* phi_t, int64_t fr,
                                                    int64_t f_r) {
  double s = 0.0;
  for (int64_t a = 0; a < Ra; ++a) {
    const double ps = psi_t[a * fl + f_l];
    if (ps == 0.0) continue;
    const double* arow = A_ij + a * a_stride;
    double t = 0.0;
    for (int64_t b = 0; b < Rb; ++b) t += arow[b] * phi_t[b * fr + f_r];
    s += ps * t;
  }
  return s;
}

}  // namespace ttx
