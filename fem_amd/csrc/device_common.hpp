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

// Sum over a block of NT threads (a multiple of 64); result valid in every
// thread.
template <int NT>
static __device__ __forceinline__ double block_sum_n(double v) {
  __shared__ double sh[NT / 64];
  __shared__ double res;
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    // pairwise: for 4 waves the order of block_sum
    double t[NT / 64];
#pragma unroll
    for (int i = 0; i < NT / 64; i++) t[i] = sh[i];
#pragma unroll
    for (int w2 = 1; w2 < NT / 64; w2 *= 2)
#pragma unroll
      for (int i = 0; i + w2 < NT / 64; i += 2 * w2) t[i] += t[i + w2];
    res = t[0];
  }
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

// ------------------------------------------------------------------------
// unit-modulus element operator in cube modes: f = E K0 u for one voxel's 24
// corner values (u[3 v + c], corner v = 4 bx + 2 by + bz), K0 = H^T Lambda H
// with H the 8-point Walsh-Hadamard transform per component and Lambda the
// 45-term mode coupling (tables.cpp)
template <typename T>
struct ModeT {
  T diag[3][8];
  T cross[3][3][8];
};

// mode coupling z = E Lambda w (mode 0 is a rigid translation: no force);
// modes sm = 4 si + 2 sj + sk per component
template <typename T>
__device__ __forceinline__ void mode_couple(const T (&w)[3][8], T E, const ModeT<T> &m, T (&z)[3][8]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    z[c][0] = 0;
#pragma unroll
    for (int sm = 1; sm < 8; sm++) {
      T acc = m.diag[c][sm] * w[c][sm];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        if (d == c) continue;
        const int bc = (sm >> (2 - c)) & 1, bd = (sm >> (2 - d)) & 1;
        if (bc == bd) continue;
        const int t = sm ^ (1 << (2 - c)) ^ (1 << (2 - d));
        acc += m.cross[c][d][sm] * w[d][t];
      }
      z[c][sm] = E * acc;
    }
  }
}

// Compact form of the same coupling.  By the symmetry of the cube under axis
// permutations the 45 nonzero coefficients take 9 values: a diagonal
// coefficient depends only on (bit of its own axis, number of other axes set)
// -> d[5], a cross coefficient (c, d) only on (bit of c, bit of the third
// axis k) -> x[4].  Passing 9 scalars instead of 96 keeps them in scalar
// registers (the full table spills there).  Indices are compile-time
// constants after unrolling.
template <typename T>
struct ModeC {
  T d[5];
  T x[4];
};
__host__ __device__ constexpr int mode_bit(int sm, int ax) { return (sm >> (2 - ax)) & 1; }
__host__ __device__ constexpr int mode_diag_index(int c, int sm) {
  // (b_c, n_other): (0,1) -> 0, (0,2) -> 1, (1,0) -> 2, (1,1) -> 3, (1,2) -> 4
  return mode_bit(sm, c) ? 2 + mode_bit(sm, (c + 1) % 3) + mode_bit(sm, (c + 2) % 3)
                         : mode_bit(sm, (c + 1) % 3) + mode_bit(sm, (c + 2) % 3) - 1;
}
__host__ __device__ constexpr int mode_cross_index(int c, int d, int sm) {
  return 2 * mode_bit(sm, c) + mode_bit(sm, 3 - c - d);
}

template <typename T>
__device__ __forceinline__ void mode_couple(const T (&w)[3][8], T E, const ModeC<T> &m, T (&z)[3][8]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    z[c][0] = 0;
#pragma unroll
    for (int sm = 1; sm < 8; sm++) {
      T acc = m.d[mode_diag_index(c, sm)] * w[c][sm];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        if (d == c) continue;
        if (mode_bit(sm, c) == mode_bit(sm, d)) continue;
        const int t = sm ^ (1 << (2 - c)) ^ (1 << (2 - d));
        acc += m.x[mode_cross_index(c, d, sm)] * w[d][t];
      }
      z[c][sm] = E * acc;
    }
  }
}

// The coupling folded into the inverse i stage, one component at a time (so
// the 24 modal forces are never all live): with z = E Lambda w,
// bottom-face forces bm[c][v] = z[c][v] - z[c][4+v] (plus the carried
// top-face forces tm of the voxel below when add_prev) and the new top-face
// forces tm[c][v] = z[c][v] + z[c][4+v].
template <typename T>
__device__ __forceinline__ void couple_split(const T (&w)[3][8], T E, const ModeC<T> &m, T (&bm)[3][4],
                                             T (&tm)[3][4], bool add_prev) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T zc[8];
    zc[0] = 0;
#pragma unroll
    for (int sm = 1; sm < 8; sm++) {
      T acc = m.d[mode_diag_index(c, sm)] * w[c][sm];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        if (d == c) continue;
        if (mode_bit(sm, c) == mode_bit(sm, d)) continue;
        const int t = sm ^ (1 << (2 - c)) ^ (1 << (2 - d));
        acc += m.x[mode_cross_index(c, d, sm)] * w[d][t];
      }
      zc[sm] = E * acc;
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const T lo = zc[v] - zc[4 + v], hi = zc[v] + zc[4 + v];
      bm[c][v] = add_prev ? lo + tm[c][v] : lo;
      tm[c][v] = hi;
    }
  }
}

template <typename T>
__device__ __forceinline__ void hadamard_voxel(const T *__restrict__ u, T E, const ModeT<T> &m,
                                               T *__restrict__ f) {
  T w[3][8];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int v = 0; v < 8; v++) w[c][v] = u[3 * v + c];
  // forward: per axis (a, b) -> (a + b, b - a)
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int mb = 4; mb >= 1; mb >>= 1)
#pragma unroll
      for (int v = 0; v < 8; v++)
        if (!(v & mb)) {
          const T a = w[c][v], b = w[c][v | mb];
          w[c][v] = a + b;
          w[c][v | mb] = b - a;
        }
  T z[3][8];
  mode_couple<T>(w, E, m, z);
  // inverse: per axis (a, b) -> (a - b, a + b)
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int mb = 4; mb >= 1; mb >>= 1)
#pragma unroll
      for (int v = 0; v < 8; v++)
        if (!(v & mb)) {
          const T a = z[c][v], b = z[c][v | mb];
          z[c][v] = a - b;
          z[c][v | mb] = a + b;
        }
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int v = 0; v < 8; v++) f[3 * v + c] = z[c][v];
}

}  // namespace femh
