// Deterministic workgroup reductions shared by the solver kernels.
#pragma once

#include <hip/hip_runtime.h>

// Takes values as operands of an empty asm statement (no instruction), so
// the compiler issues their loads before it: a kernel loads the fields it
// reads from a descriptor in one round trip ahead of a test that would
// otherwise have the loads sunk past it.
#define OP2HIP_LOADED(...) asm volatile("" ::__VA_ARGS__)

namespace op2hip {

// sum over the 64 lanes of a wavefront (every lane gets the total)
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// sum over a workgroup of NB threads, in a fixed order; the result is valid
// in thread 0.  Every thread of the workgroup must call it.
template <int NB>
__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NB / 64; ++k) t += sh[k];
  __syncthreads();
  return t;
}

}  // namespace op2hip
