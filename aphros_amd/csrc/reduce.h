// Deterministic block + grid reductions for the aphpois kernels.
//
// Every block reduces its NV partial values (sum, or max for the slots flagged
// in MaxMask) with a fixed wave64 butterfly and a fixed cross-wave order,
// writes them to partials[k * nblk + lb], and bumps the system's counter.  The
// block that arrives last re-reads all partials in block order, so the grand
// total does not depend on which block finished last.  No floating-point
// atomics are used.
#pragma once

#include "common.h"

namespace aph {

template <unsigned MaxMask, int K>
__device__ __forceinline__ double red_op(double a, double b) {
  return ((MaxMask >> K) & 1u) ? fmax(a, b) : a + b;
}

__device__ __forceinline__ double coherent_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NV, unsigned MaxMask>
__device__ __forceinline__ void wave_reduce(double (&v)[NV]) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_xor(x, off, 64);
      x = ((MaxMask >> k) & 1u) ? fmax(x, y) : x + y;
    }
    v[k] = x;
  }
}

// Reduces v over the block and, in the last block of the system, over all
// blocks.  Returns true (in every thread of that last block) with the totals
// in v; returns false elsewhere.  Must be reached by every thread of the
// block.
template <int NV, unsigned MaxMask>
__device__ bool grid_reduce(double (&v)[NV], double* partials, int nblk, int lb,
                            unsigned* counter) {
  __shared__ double red[NV][kBlock / 64];
  __shared__ int s_last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  wave_reduce<NV, MaxMask>(v);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = red[k][0];
      for (int w = 1; w < kBlock / 64; ++w) {
        t = ((MaxMask >> k) & 1u) ? fmax(t, red[k][w]) : t + red[k][w];
      }
      partials[(long long)k * nblk + lb] = t;
    }
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    s_last = (prev == (unsigned)(nblk - 1));
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const bool mx = (MaxMask >> k) & 1u;
    double t = mx ? 0.0 : 0.0;
    for (int i = threadIdx.x; i < nblk; i += kBlock) {
      const double pv = coherent_load(&partials[(long long)k * nblk + i]);
      t = mx ? fmax(t, pv) : t + pv;
    }
    v[k] = t;
  }
  wave_reduce<NV, MaxMask>(v);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][wave] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = red[k][0];
    for (int w = 1; w < kBlock / 64; ++w) {
      t = ((MaxMask >> k) & 1u) ? fmax(t, red[k][w]) : t + red[k][w];
    }
    v[k] = t;
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

}  // namespace aph
