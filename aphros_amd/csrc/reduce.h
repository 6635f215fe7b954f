// Deterministic reductions for the aphpois kernels.
//
// A reducing kernel calls block_partial(): every workgroup reduces its values
// (sum, or max for the slots flagged in MaxMask) with a fixed wave64 butterfly
// and a fixed cross-wave order and writes them to partials[k * nblk + lb] with
// plain stores.  A one-workgroup-per-system finalize kernel then calls
// reduce_partials(), which sums the partials in block order; the kernel
// boundary makes the partials visible, so no fences or atomics are needed and
// the total does not depend on scheduling.
#pragma once

#include "common.h"

namespace aph {

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

// Block-wide reduction over NT threads; the result is valid in every thread.
template <int NV, unsigned MaxMask, int NT = kBlock>
__device__ __forceinline__ void block_reduce(double (&v)[NV]) {
  __shared__ double red[NV][NT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  wave_reduce<NV, MaxMask>(v);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][wave] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = red[k][0];
    for (int w = 1; w < NT / 64; ++w) {
      t = ((MaxMask >> k) & 1u) ? fmax(t, red[k][w]) : t + red[k][w];
    }
    v[k] = t;
  }
  __syncthreads();
}

// Writes this workgroup's partial values.  Must be reached by every thread.
template <int NV, unsigned MaxMask, int NT = kBlock>
__device__ __forceinline__ void block_partial(double (&v)[NV], double* partials, int nblk,
                                              int lb) {
  block_reduce<NV, MaxMask, NT>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(long long)k * nblk + lb] = v[k];
  }
}

// Sums (or maxes) the partials of nblk workgroups in block order; result in
// every thread of the calling workgroup.
template <int NV, unsigned MaxMask>
__device__ __forceinline__ void reduce_partials(double (&v)[NV], const double* partials,
                                                int nblk) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const bool mx = (MaxMask >> k) & 1u;
    const double* pk = partials + (long long)k * nblk;
    // four independent accumulators (four loads in flight), combined in a
    // fixed order
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + 3 * kBlock < nblk; i += 4 * kBlock) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double pv = pk[i + j * kBlock];
        t[j] = mx ? fmax(t[j], pv) : t[j] + pv;
      }
    }
    for (int j = 0; i < nblk; i += kBlock, ++j) {
      const double pv = pk[i];
      t[j] = mx ? fmax(t[j], pv) : t[j] + pv;
    }
    v[k] = mx ? fmax(fmax(t[0], t[1]), fmax(t[2], t[3])) : (t[0] + t[1]) + (t[2] + t[3]);
  }
  block_reduce<NV, MaxMask>(v);
}

}  // namespace aph
