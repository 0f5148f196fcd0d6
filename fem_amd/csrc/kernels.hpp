// kernels.hpp -- launch wrappers of the gfx950 kernels (kernels.hip).
//
// Device data layout (all levels):
//   node/voxel index  (i*ny + j)*nz + k, z fastest
//   vector fields     [rhs][component][V]: RHS stride 3*V, component stride V
//   fine coefficients E[V] (Young's modulus per voxel, E = 1 stiffness K0)
//   coarse operators  27-point block stencils, S[(off*9 + 3*c + d)*V + node],
//                     off = (di+1)*9 + (dj+1)*3 + (dk+1) for the neighbour
//                     (i+di, j+dj, k+dk) with periodic wrap
// Reductions are FP64 and deterministic: each block writes a partial sum and
// the consuming kernel (or launch_finalize, for host reads) sums the partials
// in a fixed order.  No atomics or device-scope fences are involved.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace femh {

constexpr int kRHS = 6;
constexpr int kBlock = 256;
constexpr int kMaxPartials = 1024;  // partial sums per RHS per reduction
constexpr int kRedBlocks = 256;     // blocks per RHS of the vector reductions
constexpr int kSlotWidth = 18;      // scalars per slot: RHS x component

struct Dims {
  int nx, ny, nz, V;
};

// Device scalar workspace (one per solver context).  Slots are addressed by
// the enum below, each slot holds kRHS doubles.
enum ScalarSlot : int {
  S_PQ = 0,   // p . A p
  S_RZ0 = 1,  // r . z   (parity 0)
  S_RZ1 = 2,  // r . z   (parity 1)
  S_RR = 3,   // r . r
  S_BB = 4,   // b . b
  S_AUX = 5,  // power iteration / misc
  S_XR = 6,   // x . r (C_H estimate)
  S_BX = 7,   // b . x (C_H estimate)
  kNumSlots = 8
};

struct Reduce {
  double *partials;    // [kNumSlots][kSlotWidth][kMaxPartials]
  unsigned *counters;  // [kNumSlots][kSlotWidth]
  double *scalars;     // [kNumSlots][kSlotWidth]
};

// Coefficients of the mode-space element operator, staged as a kernel argument.
struct ModeCoef {
  double diag[3][8];
  double cross[3][3][8];
};

template <typename T>
struct FineOp {
  Dims g;
  const T *E;              // [V]
  const T *Ke;             // [576] unit stiffness in precision T (dense kernel)
  const ModeCoef *mode;    // host: cube-mode coupling (tiled kernel)
  int kernel;              // 0 tiled Walsh-Hadamard (default), 1 dense node gather
};

template <typename T>
struct StencilOp {
  Dims g;
  const T *S;  // [27*9][V]
};

// ---- fine level ---------------------------------------------------------
// y = A u (mode 0) or y = b - A u (mode 1) for the RHS set in mask.
// dot != 0: also reduces u . y into slot `slot`.
int dense_parts(const Dims &g);  // partials written by a dense dot-apply
int vec_parts(int V);             // partials of 3V-vector reductions
void launch_finalize(const Reduce &red, int slot, int nidx, int np, hipStream_t s);
// dense node-gather operator (diagnostic / cross-check of the tiled kernel)
template <typename T>
void launch_dense_apply(const FineOp<T> &op, const T *u, T *y, const T *b, int mode, unsigned mask,
                       const Reduce *red, int slot, hipStream_t s);
template <typename T>
void launch_fine_diag(const FineOp<T> &op, T *d, hipStream_t s);
// b[lc] = -sum_e E_e g_lc (element load of the affine field), all 6 RHS
template <typename T>
void launch_fine_rhs(const FineOp<T> &op, const double *g /*device 6x24*/, T *b, hipStream_t s);

// ---- coarse levels --------------------------------------------------------
template <typename T>
void launch_build_level1(Dims fine, Dims coarse, const T *E, const double *Mchild, T *S,
                         hipStream_t s);
template <typename T>
void launch_galerkin(Dims fine, Dims coarse, const T *Sf, T *Sc, hipStream_t s);
template <typename T>
void launch_stencil_diag(const StencilOp<T> &op, T *d, hipStream_t s);
template <typename T>
void launch_restrict(Dims fine, Dims coarse, const T *rf, T *rc, unsigned mask, hipStream_t s);
template <typename T>
void launch_prolong_add(Dims fine, Dims coarse, const T *ec, T *ef, unsigned mask, hipStream_t s);
// dense coarsest solve: x = Ainv b for each RHS (n = 3*V_coarsest)
template <typename T>
void launch_dense_solve(int n, const T *Ainv, const T *b, T *x, unsigned mask, hipStream_t s);

// ---- vector kernels -------------------------------------------------------
// Chebyshev smoothing steps (D arguments are inverse diagonals):
// first d = r/(theta D), x (+)= d; next r = rsrc - A d, d = c1 d + c2 r/D, x += d
template <typename T>
void launch_cheb_first(int V, const T *r, const T *D, T *d, T *x, double inv_theta, int x_zero,
                       unsigned mask, hipStream_t s);
template <typename T>
void launch_cheb_next(int V, const T *rsrc, T *r, const T *Ad, const T *D, T *d, T *x, double c1,
                      double c2, unsigned mask, hipStream_t s);
// PCG: alpha = rz[rzslot]/pq ; x += alpha p ; r -= alpha q ; rr -> S_RR
//      jacobi (D = inverse diagonal): z = r / D and rz -> slot rz_out
template <typename T>
void launch_pcg_update(int V, T *x, T *r, const T *p, const T *q, const T *D, T *z, int rzslot,
                       int rz_out, unsigned mask, const Reduce &red, int np_pq, hipStream_t s);
// p = z + beta p, beta = rz[new]/rz[old] (first: p = z)
template <typename T>
void launch_p_update(int V, T *p, const T *z, int rz_new, int rz_old, int first, unsigned mask,
                     const Reduce &red, hipStream_t s);
// dot products a.b -> slot (per RHS)
template <typename T>
void launch_dot(int V, const T *a, const T *b, int slot, unsigned mask, const Reduce &red,
                hipStream_t s);
// elementwise conversions / copies
template <typename Ts, typename Td>
void launch_convert(size_t n, const Ts *src, Td *dst, hipStream_t s);
// x64 += (double) d32 ; both [6][3][V]
void launch_add_f32_to_f64(size_t n, const float *d, double *x, unsigned mask, int V,
                           hipStream_t s);
// subtract per-RHS, per-component mean (uses reductions slot S_AUX)
template <typename T>
void launch_remove_mean(int V, T *x, unsigned mask, const Reduce &red, hipStream_t s);
// di = 1 / d
template <typename T>
void launch_invert(size_t n, const T *d, T *di, hipStream_t s);
// power iteration helpers (RHS slot 0 only; D arguments are inverse diagonals)
template <typename T>
void launch_fill_hash(size_t n, T *x, unsigned long long seed, hipStream_t s);
template <typename T>
void launch_scale_by_dinv_norm(int V, T *y, const T *D, const Reduce &red, hipStream_t s);
template <typename T>
void launch_normalize(int V, const T *y, T *x, const Reduce &red, hipStream_t s);

// ---- material / C_H / adjoint -------------------------------------------
// E = Emin + (E0-Emin) rho^p (interp 0) | E0 rho/(3-2rho) (1) | rho (2)
void launch_material(int V, const double *rho, double *E, int interp, double E0, double Emin,
                     double p, hipStream_t s);
// sum / non-finite count / negative count of E into S_AUX[0..2]; returns the partial count
int launch_modulus_stats(int V, const double *E, const Reduce &red, hipStream_t s);
// Energy-form C_H: C[6x6] accumulated over voxels; FP64 throughout.
// out: 21 upper-triangle sums (not yet divided by V), via partials.
void launch_ch_energy(Dims g, const double *E, const double *Ke, const double *uaff, const double *u,
                      double *partials /*[21][nblk]*/, int nblk, hipStream_t s);
void launch_sum_partials(const double *partials, int nsum, int nblk, double *out, hipStream_t s);
// dJ/drho_e = dE/drho(rho_e) * sum_p Gs[p] W_p(e) ; W from the fields
void launch_gradient(Dims g, const double *rho, const double *Ke, const double *uaff,
                     const double *u, const double *Gs /*device 21*/, int interp, double E0,
                     double Emin, double p, int energy_only, double *out, hipStream_t s);

void reduce_init(Reduce &red, hipStream_t s);

}  // namespace femh
