// Deterministic block / grid reductions of FP64 partial sums (device side).
#pragma once

#include <hip/hip_runtime.h>

namespace tpf {

constexpr int kBlock = 256;

// FP32 stencil sweeps: each workgroup takes one contiguous chunk of the index
// range, and the chunks are numbered XCD by XCD (workgroups are dispatched
// to the 8 XCDs round-robin), so an XCD sweeps one contiguous slab and the
// +-1 row / plane neighbours a workgroup reads were fetched into the same
// L2 by its neighbouring workgroups.  Measured on the multigrid levels
// (257^3 W-cycle FCG 32.7 -> 28.0 ms); the FP64 CG kernels stay on
// grid-stride loops, which measured faster there (one wide sweep front).
__device__ inline void xcd_chunk(int64_t total, int64_t& begin, int64_t& end) {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned lb = nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
  const int64_t chunk = (total + nb - 1) / nb;
  begin = (int64_t)lb * chunk;
  end = begin + chunk < total ? begin + chunk : total;
}
#define CHUNK_LOOP(idx, total)                                   \
  int64_t idx##_b, idx##_e;                                      \
  xcd_chunk((total), idx##_b, idx##_e);                          \
  for (int64_t idx = idx##_b + threadIdx.x; idx < idx##_e; idx += kBlock)

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

// Writes this block's K partial sums to partials[blockIdx.x * K + k].  The
// grid totals are formed by a following single-block kernel (finish_sum):
// kernel boundaries order the writes, which is far cheaper on a multi-XCD
// device than a per-block agent-scope release and atomic.
template <int K>
__device__ inline void block_partials(double (&v)[K], double* partials, double* sh) {
  block_sum<K>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) partials[blockIdx.x * K + k] = v[k];
}

// Single-block reduction of `nblocks` partial rows; thread 0 gets the totals.
// Fixed thread assignment and tree order: deterministic.
template <int K>
__device__ inline void finish_sum(double (&v)[K], const double* partials, int nblocks,
                                  double* sh) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += partials[b * K + k];
  block_sum<K>(v, sh);
}

}  // namespace tpf
