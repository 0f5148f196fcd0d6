// tables.hpp -- host-side element tables for the voxel engine.
//
// The unit-voxel trilinear hex stiffness is built here from its closed form in
// the Walsh-Hadamard ("cube mode") basis rather than by quadrature:
//
//   K0 = H^T Lambda H,   H[s][b] = prod_axis (2 b_axis - 1)^{s_axis}
//
// with s, b in {0,1}^3 (local index 4x + 2y + z, z fastest -- the corner order
// of femlib/src/ElementHex.cpp:24-33) applied per displacement component.
// Lambda couples component c of mode s with itself and with component d of
// mode s ^ e_c ^ e_d only; its 45 non-zeros follow from the exact 2-point
// Gauss integrals of products of the 1D modes (1/2, x-1/2):
//   int psi_0^2 = 1/4, int psi_1^2 = 1/12, int psi_0 = 1/2, int psi_1 = 0.
// This is the same matrix femlib/src/MaterialQuad.cpp:68-105 integrates
// (mu (dF + dF^T) + lambda tr(dF) I, femlib/src/StrainLin.cpp:29-37) at h = 1.
#pragma once
#include <array>
#include <vector>

namespace femh {

struct ModeCoupling {
  // diag[c][s]: Lambda[(s,c),(s,c)]
  double diag[3][8];
  // cross[c][d][s]: Lambda[(s,c),(s^e_c^e_d, d)] for d != c (0 when absent)
  double cross[3][3][8];
};

struct ElementTables {
  double nu = 0.3, lambda = 0, mu = 0;
  ModeCoupling L;                      // Hadamard-space coupling (E = 1)
  std::array<double, 576> Ke{};        // K0 (E = 1, h = 1), row-major, dof 3v+c
  std::array<double, 6 * 24> g{};      // K0 * u_affine(load case), per case
  std::array<double, 6 * 24> u_aff{};  // affine local field per load case
  std::vector<double> Mchild;          // 8 x 576: P_f^T K0 P_f (level-1 Galerkin)
  // Coupling of the coarse-strain (force-driven) form of C_H: the same
  // element integral with the engineering stiffness applied to tensor shear
  // strains on both sides, i.e. T C0 T with T = diag(1,1,1,1/2,1/2,1/2) --
  // the stiffness of a cubic material c11 = lambda + 2 mu, c12 = lambda,
  // c44 = mu / 4 (unit modulus)
  ModeCoupling LT;
};

// Hadamard-space coupling of a cubic material (c11, c12, c44) on a unit
// voxel; the isotropic one is c11 = lambda + 2 mu, c12 = lambda, c44 = mu
ModeCoupling cubic_coupling(double c11, double c12, double c44);

void lame(double E, double nu, double *lambda, double *mu);
void macro_strain(int lc, double eps[9]);
ElementTables build_tables(double nu);

}  // namespace femh
