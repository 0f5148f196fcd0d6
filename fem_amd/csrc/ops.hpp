// ops.hpp -- fused operator kernels (ops.hip): K u with a staging transform
// applied to the input and a Chebyshev/PCG epilogue applied to the output,
// so that one smoothing step is one kernel.
//
// Staging (value of the input at every node the stencil touches):
//   ST_PLAIN    u
//   ST_DSCALE   u * theta_inv / D              (first Chebyshev step from x = 0)
//   ST_PROLONG  u + P xc                       (coarse correction added on the fly)
//   ST_PUPD     u + beta u2                    (PCG direction p = z + beta p_old,
//                                               beta per RHS from a device array)
// (D enters as its precomputed inverse: no divisions in the kernels)
// Epilogue at each owned node, with Au the operator result and u_own the
// staged value there:
//   EP_AX       y = Au            (optional FP64 partial of u . Au; with pout,
//                                  pout = u_own: the new PCG direction; with
//                                  ydinv, y = Au / D: a power-iteration step)
//   EP_RES      y = b - Au            (optional FP64 partials of r.r, b.u, u.r)
//   EP_CHEB0    r = b - Au ; dout = r theta_inv / D (not with `last`) ; x = u_own
//               with fin1 (a degree-1 smoothing that ends the V-cycle): only
//               x = u_own + r theta_inv / D, and with `dot` the FP64 partial
//               of b . x
//   EP_STEP     r = rprev - Au ; dn = c1 u_own + c2 r / D ; dout = dn ;
//               (with `dot`, fine level: the FP64 partial of b . x_new)
//               x = xprev + u_own ; with `last`: x += dn, and r and dout are
//               not written (nothing reads them after the last step)
// A kernel never writes the vector it stages (neighbours read it), so x/xn,
// d/d2 are ping-ponged by the caller.  Chebyshev coefficients are per sample
// (each sample's smoother has its own spectrum estimate).
#pragma once
#include "kernels.hpp"

namespace femh {

enum Stage : int { ST_PLAIN = 0, ST_DSCALE = 1, ST_PROLONG = 2, ST_PUPD = 3 };
enum Epi : int { EP_AX = 0, EP_RES = 1, EP_CHEB0 = 2, EP_STEP = 3 };
constexpr int kMaxCheb = 16;

template <typename T>
struct OpArgs {
  const T *u = nullptr;      // staged vector [R][3][V]
  const T *u2 = nullptr;     // ST_PUPD: previous direction [R][3][V]
  const double *beta = nullptr;  // ST_PUPD: per-RHS beta [R]
  T *pout = nullptr;         // EP_AX: staged value written back (new direction)
  int ydinv = 0;             // EP_AX: y = Dinv * Au
  int fin1 = 0;              // EP_CHEB0: write only the smoothed x (see above)
  const T *xc = nullptr;     // ST_PROLONG: coarse vector [R][3][Vc]
  Dims gc{};                 // ST_PROLONG: coarse dims
  const T *Dinv = nullptr;   // inverse diagonal [B][3][V]
  T *y = nullptr;
  const T *b = nullptr;
  T *r = nullptr;
  const T *rprev = nullptr;
  T *dout = nullptr;
  const T *xprev = nullptr;
  T *x = nullptr;
  int last = 0;
  const int *act = nullptr;  // [R] activity flags
  Reduce red{};
  int slot = 0;
  int dot = 0;               // EP_RES: r.r -> slot, b.x -> slot2, x.r -> slot3
  int slot2 = 0, slot3 = 0;
  // Chebyshev coefficients of this launch, device memory [3][kMaxB]: 1/theta,
  // c1, c2 per sample (null where the variant uses none).  In device memory,
  // not in the arguments, so that a recorded launch sequence stays valid when
  // a new solve brings new spectra.
  const double *cheb = nullptr;
  // outputs stored non-temporally (set by the fine-level launcher for
  // vectors too large to stay in the caches until the next launch reads them)
  int nt = 0;
};

// fine level (tiled cube-mode voxel kernel).  shape: launch shape, bits 0-7
// node planes per workgroup and bits 8-15 the tile variant (TJ x TK owned
// node columns per workgroup: 1 = 16 x 16, 2 = 8 x 32, 3 = 16 x 32); 0 in a
// field is the built-in choice (tiled_default_tile, planes from the grid).
// tiled_parts: partial count of a dot launch with that shape; tiled_slabs:
// the planes per workgroup actually used (both depend on the element size
// sz, 4 or 8, for a one-cell launch).  ev0/ev1 (optional): events the
// runtime records at the kernel's own start and end (kernel timing).
constexpr int kNumTileShapes = 4;
inline int tiled_shape_code(int tile, int planes) { return (tile << 8) | planes; }
int tiled_default_tile(const Dims &g, int R);  // the tile variant of the built-in shape
template <typename T>
void launch_tiled(int st, int ep, const Dims &g, int R, const T *E, const ModeCoef &mc,
                  const OpArgs<T> &a, hipStream_t s, int shape = 0, hipEvent_t ev0 = nullptr,
                  hipEvent_t ev1 = nullptr);
int tiled_parts(const Dims &g, int R, int shape, int sz);
int tiled_slabs(const Dims &g, int R, int shape, int sz);
// launch shape for a launch with only Ract of its R right-hand sides active
// (the power iteration: one per cell) and no fused dot: shape's tile, the
// chunk length from the rounds of the active workgroups (shape returned as
// it is when its planes are pinned)
int tiled_sparse_shape(const Dims &g, int R, int Ract, int shape, int sz);

// coarse level (27-point block stencil, one RHS per thread)
template <typename T>
void launch_stencil(int st, int ep, const Dims &g, int R, const T *S, const OpArgs<T> &a,
                    hipStream_t s);

// x = xn + d (closing a degree-1 post-smoothing)
template <typename T>
void launch_add2(int V, int R, const T *xn, const T *d, T *x, const int *act, hipStream_t s);

}  // namespace femh
