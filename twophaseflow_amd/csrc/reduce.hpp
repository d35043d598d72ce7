// Deterministic block / grid reductions of FP64 partial sums (device side).
#pragma once

#include <hip/hip_runtime.h>

namespace tpf {

constexpr int kBlock = 256;

// Sums K doubles over a block of kBlock threads; thread 0 gets the totals.
template <int K>
__device__ inline void block_sum(double (&v)[K], double* sh) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double x = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[wid * K + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += sh[w * K + k];
      v[k] = s;
    }
  }
}

// Writes this block's partials; returns true in exactly one (the last) block,
// whose thread 0 then holds the grid totals in v.  The counter is an
// acq_rel agent-scope atomic, so after it the last block reads the partials
// with plain loads; it resets the counter for the next launch.
template <int K>
__device__ inline bool grid_sum(double (&v)[K], double* partials, unsigned* counter, double* sh) {
  block_sum<K>(v, sh);
  __shared__ bool last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[blockIdx.x * K + k] = v[k];
    const unsigned t =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  // fixed assignment of blocks to threads + fixed tree => deterministic
  for (unsigned b = threadIdx.x; b < gridDim.x; b += kBlock)
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += partials[b * K + k];
  block_sum<K>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = acc[k];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

}  // namespace tpf
