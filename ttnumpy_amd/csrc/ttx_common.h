// Shared helpers of the ttx device library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ttx.h"

namespace ttx {

constexpr int kWave = 64;

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

}  // namespace ttx
