// kernels.hpp -- launch wrappers of the gfx950 kernels (kernels.hip, ops.hip).
//
// Batching: a context solves B samples of one grid size together.  Every
// vector field holds R = 6B right-hand sides (sample s, load case lc at
// q = 6 s + lc); per-sample operator data (moduli, stencils, diagonals, dense
// coarse inverses) is stored sample after sample.  Kernels take the RHS index
// from their grid and skip a RHS whose entry in the device activity array
// `act[q]` is zero.
//
// Device data layout (all levels):
//   node/voxel index  (i*ny + j)*nz + k, z fastest
//   vector fields     [q][component][V]: RHS stride 3*V, component stride V
//   fine moduli       E[s][V] (unit-modulus stiffness K0 per voxel)
//   coarse operators  27-point block stencils, 243*V values per sample: block
//                     off = (di+1)*9 + (dj+1)*3 + (dk+1) couples a node to its
//                     neighbour (i+di, j+dj, k+dk) (periodic wrap); its entry
//                     q = 3*c + d of node n sits at stencil_index(off, q, V, n)
//   diagonals         D[s][c][V] and its inverse Dinv (what the kernels use)
// Reductions are FP64 and deterministic: each block writes a partial sum and
// the consuming kernel (or launch_finalize, for host reads) sums the partials
// in a fixed order.  No atomics or device-scope fences are involved.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace femh {

// Offset of entry q (row-major in the 3x3 block) of stencil block `off` at
// node n, in one sample's stencil of V nodes.  Each block is 9*V values:
// entries 0-3 of every node as [V][4], entries 4-7 as [V][4], entry 8 as
// [V], so a node's block is two 16-byte loads and one scalar load.
__host__ __device__ inline size_t stencil_index(int off, int q, int V, int n) {
  const size_t b = (size_t)off * 9 * (size_t)V;
  if (q < 4) return b + (size_t)n * 4 + (size_t)q;
  if (q < 8) return b + 4 * (size_t)V + (size_t)n * 4 + (size_t)(q - 4);
  return b + 8 * (size_t)V + (size_t)n;
}

constexpr int kRHS = 6;             // load cases per sample
constexpr int kMaxB = 16;           // samples per context
constexpr int kMaxR = kRHS * kMaxB; // right-hand sides per context
constexpr int kBlock = 256;
constexpr int kMaxPartials = 1024;  // partial sums per RHS per reduction
constexpr int kRedBlocks = 256;     // blocks per RHS of the vector reductions
constexpr int kSlotWidth = 3 * kMaxR;  // scalars per slot: RHS x component

struct Dims {
  int nx, ny, nz, V;
};

enum ScalarSlot : int {
  S_PQ = 0,   // p . A p
  S_RZ0 = 1,  // r . z   (parity 0)
  S_RZ1 = 2,  // r . z   (parity 1)
  S_RR = 3,   // r . r
  S_BB = 4,   // b . b
  S_AUX = 5,  // power iteration / means / moduli statistics
  S_XR = 6,   // x . r (C_H estimate)
  S_BX = 7,   // b . x (C_H estimate)
  kNumSlots = 8
};

struct Reduce {
  double *partials;    // [kNumSlots][kSlotWidth][kMaxPartials]
  unsigned *counters;  // [kNumSlots][kSlotWidth]
  double *scalars;     // [kNumSlots][kSlotWidth]
};

// Coefficients of the mode-space element operator (host side).
struct ModeCoef {
  double diag[3][8];
  double cross[3][3][8];
};

template <typename T>
struct FineOp {
  Dims g;
  const T *E;   // [B][V]
  const T *Ke;  // [576] unit stiffness in precision T (dense kernel)
};

template <typename T>
struct StencilOp {
  Dims g;
  const T *S;  // [B][243][V]
};

int dense_parts(const Dims &g);  // partials written by a dense dot-apply
int vec_parts(int V);             // partials of 3V-vector reductions
void launch_finalize(const Reduce &red, int slot, int nidx, int np, hipStream_t s);

// ---- fine level ---------------------------------------------------------
// dense node-gather operator (diagnostic / cross-check of the tiled kernel):
// y = A u (mode 0) or b - A u (mode 1); red != null: u . y into `slot`
template <typename T>
void launch_dense_apply(const FineOp<T> &op, int R, const T *u, T *y, const T *b, int mode,
                        const int *act, const Reduce *red, int slot, hipStream_t s);
template <typename T>
void launch_fine_diag(const FineOp<T> &op, int B, T *d, hipStream_t s);
// b[q] = -sum_e E_e g_lc (element load of the affine field), with b.b of
// every RHS into slot (vec_parts(V) partials) and, when b32 is given, an FP32
// copy of b
template <typename T>
void launch_fine_rhs(const FineOp<T> &op, int R, const double *g /*device 6x24*/, T *b, float *b32,
                     const Reduce &red, int slot, hipStream_t s);

// ---- coarse levels --------------------------------------------------------
template <typename T>
void launch_build_level1(Dims fine, Dims coarse, int B, const T *E, const double *Mchild, T *S,
                         hipStream_t s);
// coarse stencil P^T A P by three one-axis passes; t1 holds 243 B V/2 and t2
// 243 B V/4 values of the fine level's V
template <typename T>
void launch_galerkin(Dims fine, Dims coarse, int B, const T *Sf, T *Sc, T *t1, T *t2, hipStream_t s);
template <typename T>
void launch_stencil_diag(const StencilOp<T> &op, int B, T *d, hipStream_t s);
// planes: coarse planes per thread (0: the built-in choice; an autotuned
// value otherwise -- the results do not depend on it)
template <typename T>
void launch_restrict(Dims fine, Dims coarse, int R, const T *rf, T *rc, const int *act,
                     hipStream_t s, int planes = 0);
// ef += P ec; with out, out = ef + P ec (ef is then only read)
template <typename T>
void launch_prolong_add(Dims fine, Dims coarse, int R, const T *ec, T *ef, const int *act,
                        hipStream_t s, T *out = nullptr, int planes = 0);
// dense coarsest solve: x = Ainv[s] b for each RHS (n = 3*V_coarsest)
template <typename T>
void launch_dense_solve(int n, int R, const T *Ainv, const T *b, T *x, const int *act,
                        hipStream_t s);

// ---- vector kernels (D arguments are inverse diagonals, per sample) ------
// d = r theta_inv / D ; x = d (x_zero) or x += d
template <typename T>
void launch_cheb_first(int V, int R, const T *r, const T *D, T *d, T *x, double inv_theta,
                       int x_zero, const int *act, hipStream_t s);
// PCG: alpha = rz[rzslot]/pq ; x += alpha p ; r -= alpha q ; rr -> S_RR
//      jacobi (D != null): z = r / D and rz -> slot rz_out (vec_parts(V) partials)
// np_pq / np_rz: partial counts of the slots read (set by their producers)
// alpha_out (may be null): alpha of every active RHS (the Lanczos record,
// [iteration][kMaxR], row cnt[r] = the iterations RHS r has completed)
template <typename T>
void launch_pcg_update(int V, int R, T *x, T *r, const T *p, const T *q, const T *D, T *z,
                       int rzslot, int rz_out, const int *act, const Reduce &red, int np_pq,
                       int np_rz, hipStream_t s, double *alpha_out = nullptr, const int *cnt = nullptr);
// p = z + beta p, beta = rz[new]/rz[old] (first: p = z); beta_out as
// alpha_out, row cnt[r] - 1 (after the iteration's convergence check)
template <typename T>
void launch_p_update(int V, int R, T *p, const T *z, int rz_new, int rz_old, int first,
                     const int *act, const Reduce &red, int np_rz, hipStream_t s,
                     double *beta_out = nullptr, const int *cnt = nullptr);
// beta[r] = rz[new]/rz[old] per active RHS (first: 0), for the fused
// direction update of the fine operator (ST_PUPD); beta_out as for p_update
void launch_beta(int R, int rz_new, int rz_old, int first, const int *act, const Reduce &red, int np_rz,
                 double *beta, hipStream_t s, double *beta_out = nullptr, const int *cnt = nullptr);
// PCG convergence check on the device (after the update of an iteration):
// for every active RHS, cnt += 1 and relres = sqrt(r.r / bb) from the np
// partials of S_RR; relres <= tol clears act (state 1), a non-finite r.r too
// (state 2)
void launch_pcg_check(int R, int *act, const Reduce &red, int np, const double *bb, double tol, int *cnt,
                      double *relres, int *state, hipStream_t s);
// cubic structure: load case dst from load case src by an axis swap of the
// cube grid (swap 0: x <-> y, 1: x <-> z), points and displacement components
// alike, for the samples s with bit s of mask set
void launch_permute_case(int n, int nb, double *x, int src, int dst, int swap, unsigned mask, hipStream_t s);
// dot products a.b -> slot (per RHS)
template <typename T>
void launch_dot(int V, int R, const T *a, const T *b, int slot, const int *act, const Reduce &red,
                hipStream_t s);
template <typename Ts, typename Td>
void launch_convert(size_t n, const Ts *src, Td *dst, hipStream_t s);
template <typename T>
void launch_invert(size_t n, const T *d, T *di, hipStream_t s);
// x[t] = v for t < n (clears memory with a kernel, not a memset node, so that
// the sequence stays valid under stream capture)
template <typename T>
void launch_fill(size_t n, T *x, T v, hipStream_t s);
// x[q] = 0 for the right-hand sides q flagged in act
void launch_zero_rhs(int V, int R, double *x, const int *act, hipStream_t s);
// x64 += (double) d32 ; both [R][3][V]
void launch_add_f32_to_f64(int V, int R, const float *d, double *x, const int *act, hipStream_t s);
// subtract each RHS/component mean (slot S_AUX)
template <typename T>
void launch_remove_mean(int V, int R, T *x, const int *act, const Reduce &red, hipStream_t s);
// x[t] = hash(seed, t mod period) in [-1/2, 1/2)
template <typename T>
void launch_fill_hash(size_t n, size_t period, T *x, unsigned long long seed, hipStream_t s);
// y *= Dinv and ||y||^2 -> S_AUX[q] (the D^-1 of a power-iteration step on
// the dense fine-apply path)
template <typename T>
void launch_scale_by_dinv_norm(int V, int R, T *y, const T *Dinv, const int *act,
                               const Reduce &red, hipStream_t s);

// ---- material / C_H / adjoint -------------------------------------------
// E = Emin + (E0-Emin) rho^p (interp 0) | E0 rho/(3-2rho) (1) | rho (2)
void launch_material(size_t n, const double *rho, double *E, int interp, double E0, double Emin,
                     double p, hipStream_t s);
// per sample: sum / non-finite count / negative count of E into S_AUX[3s..3s+2];
// cube_n > 0 (cube grid of that edge): the count of voxels not invariant under
// the x <-> y and x <-> z swaps into S_AUX[3B + s]
int launch_modulus_stats(int V, int B, const double *E, const Reduce &red, hipStream_t s, int cube_n = 0);
// Energy-form C_H per sample: 21 upper-triangle sums (not yet / V) into out[s][21]
void launch_ch_energy(Dims g, int B, const double *E, const ModeCoef &m, const double *uaff,
                      const double *u, double *partials /*[B][21][nblk]*/, int nblk, double *out,
                      hipStream_t s);
// thermal stress sums per sample: out[s][q] = sum_e alpha_e E_e mean div(chi_q)
// over element e (not yet scaled by 1/((1 - 2 nu) V)); partials [B][6][nblk]
void launch_thermal(Dims g, int B, const double *E, const double *alpha, const double *uaff, const double *u,
                    double *partials, int nblk, double *out, hipStream_t s);
// dJ/drho_e = dE/drho(rho_e) * sum_p Gs[s][p] W_p(e) / V ; energy_only: W (21/voxel)
void launch_gradient(Dims g, int B, const double *rho, const ModeCoef &m, const double *uaff,
                     const double *u, const double *Gs /*device [B][21]*/, int interp, double E0,
                     double Emin, double p, int energy_only, double *out, hipStream_t s);

// Projected-gradient trial point of n cells of V voxels: per cell, h > 0:
// out = clamp(x - h g, lb, ub); h < 0: out = x; h == 0: out unchanged
// (h: device [n]); partials [n][nblk] receive the block sums of out.
void launch_descent_trial(int n, long long V, const double *x, const double *g, const double *h, double lb,
                          double ub, double *out, double *partials, int nblk, hipStream_t s);
// g[s] += shift[s] (shift: device [n] or null); partials [n][nblk] receive the
// block maxima of |g[s]|
void launch_shift_absmax(int n, long long V, double *g, const double *shift, double *partials, int nblk,
                         hipStream_t s);

void reduce_init(Reduce &red, hipStream_t s);

}  // namespace femh
