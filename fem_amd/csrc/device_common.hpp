// device_common.hpp -- block reductions and index helpers shared by the
// kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace femh {

// ------------------------------------------------------------------------
// reductions
static __device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a 256-thread block; result valid in every thread.
static __device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kBlock / 64];
  __shared__ double res;
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) res = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return res;
}

// Block partial of a reduction: thread 0 stores the block's sum at
// partials[slot][idx][pidx].  No atomics and no device-scope fences: the
// partials are consumed by a later kernel (kernel boundaries order them).
static __device__ __forceinline__ void block_partial(const Reduce &red, int slot, int idx, int pidx,
                                              double bsum) {
  if (threadIdx.x == 0) red.partials[((size_t)slot * kSlotWidth + idx) * kMaxPartials + pidx] = bsum;
}

// Sum of the first np partials of (slot, idx) in a fixed order; every thread
// of the block receives the value.  All threads must call.
static __device__ __forceinline__ double sum_parts(const Reduce &red, int slot, int idx, int np) {
  const double *pp = red.partials + ((size_t)slot * kSlotWidth + idx) * kMaxPartials;
  double s = 0;
  for (int i = threadIdx.x; i < np; i += kBlock) s += pp[i];
  return block_sum(s);
}

static inline int grid_for(int V) {
  int b = (V + kBlock - 1) / kBlock;
  return b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks;
}

static __device__ __forceinline__ void node_ijk(const Dims &g, int n, int &i, int &j, int &k) {
  k = n % g.nz;
  int t = n / g.nz;
  j = t % g.ny;
  i = t / g.ny;
}

}  // namespace femh
