// ops.hpp -- fused operator kernels (ops.hip): K u with a staging transform
// applied to the input and a Chebyshev/PCG epilogue applied to the output,
// so that one smoothing step is one kernel.
//
// Staging (value of the input at every node the stencil touches):
//   ST_PLAIN    u
//   ST_DSCALE   u * theta_inv / D              (first Chebyshev step from x = 0)
// (D enters as its precomputed inverse: no divisions in the kernels)
//   ST_PROLONG  u + P xc                       (coarse correction added on the fly)
// Epilogue at each owned node, with Au the operator result and u_own the
// staged value there:
//   EP_AX       y = Au            (optional FP64 partial of u . Au)
//   EP_RES      y = b - Au
//   EP_CHEB0    r = b - Au ; dout = r theta_inv / D ; x = u_own
//   EP_STEP     r = rprev - Au ; dn = c1 u_own + c2 r / D ; dout = dn ;
//               x = xprev + u_own (+ dn when `last`)
// A kernel never writes the vector it stages (neighbours read it), so x/xn,
// d/d2 are ping-ponged by the caller.
#pragma once
#include "kernels.hpp"

namespace femh {

enum Stage : int { ST_PLAIN = 0, ST_DSCALE = 1, ST_PROLONG = 2 };
enum Epi : int { EP_AX = 0, EP_RES = 1, EP_CHEB0 = 2, EP_STEP = 3 };

template <typename T>
struct OpArgs {
  const T *u = nullptr;      // staged vector [6][3][V]
  const T *xc = nullptr;     // ST_PROLONG: coarse vector [6][3][Vc]
  Dims gc{};                 // ST_PROLONG: coarse dims
  const T *Dinv = nullptr;   // inverse diagonal [3][V]
  T theta_inv = 1;
  T *y = nullptr;
  const T *b = nullptr;
  T *r = nullptr;
  const T *rprev = nullptr;
  T *dout = nullptr;
  const T *xprev = nullptr;
  T *x = nullptr;
  T c1 = 0, c2 = 0;
  int last = 0;
  unsigned mask = 0;
  Reduce red{};
  int slot = 0;
  int dot = 0;
};

// fine level (tiled cube-mode voxel kernel); partial count of a dot launch
template <typename T>
void launch_tiled(int st, int ep, const Dims &g, const T *E, const ModeCoef &mc, const OpArgs<T> &a,
                  hipStream_t s);
int tiled_parts(const Dims &g);

// coarse level (27-point block stencil, all RHS per thread)
template <typename T>
void launch_stencil(int st, int ep, const Dims &g, const T *S, const OpArgs<T> &a, hipStream_t s);

// x = xn + d (closing a degree-1 post-smoothing)
template <typename T>
void launch_add2(int V, const T *xn, const T *d, T *x, unsigned mask, hipStream_t s);

// Whole V-cycle below (and including) a small level, run by one workgroup.
constexpr int kMaxCheb = 16;
// Levels with at most this many nodes join the single-workgroup cycle.  0:
// only the coarsest (dense) solve runs there -- measured faster than fusing
// the 8^3 / 4^3 levels, whose single-workgroup sweeps are latency bound.
constexpr int kSmallMaxV = 0;
constexpr int kSmallThreads = 512;

template <typename T>
struct SmallLevel {
  Dims g;
  const T *S;   // stencil [243][V] (unused on the coarsest dense level)
  const T *Dinv;
  T *x, *b, *r, *d, *d2, *xn;  // [6][3][V]
  T theta_inv;
  T c1[kMaxCheb], c2[kMaxCheb];  // Chebyshev step coefficients
  const T *Ainv;  // coarsest only: dense inverse, n = 3V
};

template <typename T>
void launch_small_cycle(const SmallLevel<T> *d_levels, int nlev, int nu1, int nu2, unsigned mask,
                        hipStream_t s);

}  // namespace femh
