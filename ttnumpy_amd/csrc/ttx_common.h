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

inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? TTX_OK : TTX_EHIP;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
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
