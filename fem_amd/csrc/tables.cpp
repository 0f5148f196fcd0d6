// tables.cpp -- closed-form element tables (see tables.hpp for the derivation).
#include "tables.hpp"

#include <cstring>

namespace femh {

void lame(double E, double nu, double *lambda, double *mu) {
  // Utilities/src/cfgMaterialUtilities.cpp:2945-2955
  *lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  *mu = E / (2.0 * (1.0 + nu));
}

void macro_strain(int lc, double eps[9]) {
  std::memset(eps, 0, 9 * sizeof(double));
  switch (lc) {
    case 0: eps[0] = 1; break;
    case 1: eps[4] = 1; break;
    case 2: eps[8] = 1; break;
    case 3: eps[5] = eps[7] = 0.5; break;  // yz, engineering shear 1
    case 4: eps[2] = eps[6] = 0.5; break;  // xz
    case 5: eps[1] = eps[3] = 0.5; break;  // xy
    default: break;
  }
}

static inline int bit(int s, int ax) { return (s >> (2 - ax)) & 1; }

// Same-component block: c11 int d_c N d_c N + c44 sum_{i != c} int d_i N d_i N;
// cross block (c, d): c12 int d_c N_a d_d N_b + c44 int d_d N_a d_c N_b.  For
// the isotropic material (c11, c12, c44) = (lambda + 2 mu, lambda, mu) the
// diagonal reads (lambda + mu) b_c W(c) + mu sum_i b_i W(i).
ModeCoupling cubic_coupling(double c11, double c12, double c44) {
  ModeCoupling L;
  const double w[2] = {0.25, 1.0 / 12.0};
  std::memset(&L, 0, sizeof(L));
  for (int s = 0; s < 8; s++) {
    for (int c = 0; c < 3; c++) {
      auto W = [&](int i) {  // product of w over the two axes other than i
        double p = 1;
        for (int k = 0; k < 3; k++)
          if (k != i) p *= w[bit(s, k)];
        return p;
      };
      double D = c11 * bit(s, c) * W(c);
      for (int i = 0; i < 3; i++)
        if (i != c) D += c44 * bit(s, i) * W(i);
      L.diag[c][s] = D;
      for (int d = 0; d < 3; d++) {
        if (d == c) continue;
        int k = 3 - c - d;
        double v = 0;
        if (bit(s, c) == 1 && bit(s, d) == 0) v = c12 * 0.25 * w[bit(s, k)];
        if (bit(s, c) == 0 && bit(s, d) == 1) v = c44 * 0.25 * w[bit(s, k)];
        L.cross[c][d][s] = v;
      }
    }
  }
  return L;
}

ElementTables build_tables(double nu) {
  ElementTables t;
  t.nu = nu;
  lame(1.0, nu, &t.lambda, &t.mu);
  const double lam = t.lambda, mu = t.mu;
  // (lambda + mu) b_c W(c) + mu sum_i b_i W(i), as the isotropic case of the
  // cubic coupling but summed in the order used since the first tables
  {
    const double w[2] = {0.25, 1.0 / 12.0};
    std::memset(&t.L, 0, sizeof(t.L));
    for (int s = 0; s < 8; s++) {
      for (int c = 0; c < 3; c++) {
        auto W = [&](int i) {
          double p = 1;
          for (int k = 0; k < 3; k++)
            if (k != i) p *= w[bit(s, k)];
          return p;
        };
        double D = (lam + mu) * bit(s, c) * W(c);
        for (int i = 0; i < 3; i++) D += mu * bit(s, i) * W(i);
        t.L.diag[c][s] = D;
        for (int d = 0; d < 3; d++) {
          if (d == c) continue;
          int k = 3 - c - d;
          double v = 0;
          if (bit(s, c) == 1 && bit(s, d) == 0) v = lam * 0.25 * w[bit(s, k)];
          if (bit(s, c) == 0 && bit(s, d) == 1) v = mu * 0.25 * w[bit(s, k)];
          t.L.cross[c][d][s] = v;
        }
      }
    }
  }
  t.LT = cubic_coupling(lam + 2.0 * mu, lam, 0.25 * mu);
  // K0 = H^T Lambda H
  double H[8][8];
  for (int s = 0; s < 8; s++)
    for (int b = 0; b < 8; b++) {
      double v = 1;
      for (int ax = 0; ax < 3; ax++)
        if (bit(s, ax)) v *= (2 * bit(b, ax) - 1);
      H[s][b] = v;
    }
  double Lam[24][24];
  std::memset(Lam, 0, sizeof(Lam));
  for (int s = 0; s < 8; s++)
    for (int c = 0; c < 3; c++) {
      Lam[3 * s + c][3 * s + c] = t.L.diag[c][s];
      for (int d = 0; d < 3; d++) {
        if (d == c) continue;
        int tt = s ^ (1 << (2 - c)) ^ (1 << (2 - d));
        Lam[3 * s + c][3 * tt + d] += t.L.cross[c][d][s];
      }
    }
  for (int a = 0; a < 8; a++)
    for (int c = 0; c < 3; c++)
      for (int b = 0; b < 8; b++)
        for (int d = 0; d < 3; d++) {
          double acc = 0;
          for (int s = 0; s < 8; s++)
            for (int u = 0; u < 8; u++) acc += H[s][a] * Lam[3 * s + c][3 * u + d] * H[u][b];
          t.Ke[(3 * a + c) * 24 + 3 * b + d] = acc;
        }
  // affine local fields and K0 * u_affine
  for (int lc = 0; lc < 6; lc++) {
    double eps[9];
    macro_strain(lc, eps);
    for (int v = 0; v < 8; v++) {
      double x[3] = {double((v >> 2) & 1), double((v >> 1) & 1), double(v & 1)};
      for (int c = 0; c < 3; c++)
        t.u_aff[lc * 24 + 3 * v + c] = eps[3 * c] * x[0] + eps[3 * c + 1] * x[1] + eps[3 * c + 2] * x[2];
    }
    for (int r = 0; r < 24; r++) {
      double acc = 0;
      for (int q = 0; q < 24; q++) acc += t.Ke[r * 24 + q] * t.u_aff[lc * 24 + q];
      t.g[lc * 24 + r] = acc;
    }
  }
  // level-1 Galerkin child matrices: P_f maps the 8 corners of a coarse
  // (2x2x2-voxel) element to the corners of its child f by trilinear
  // interpolation; M_f = P_f^T K0 P_f.
  t.Mchild.assign(8 * 576, 0.0);
  for (int f = 0; f < 8; f++) {
    double P[24][24];
    std::memset(P, 0, sizeof(P));
    for (int a = 0; a < 8; a++)
      for (int A = 0; A < 8; A++) {
        double wgt = 1;
        for (int ax = 0; ax < 3; ax++) {
          double x = 0.5 * (bit(f, ax) + bit(a, ax));
          wgt *= bit(A, ax) ? x : 1.0 - x;
        }
        for (int c = 0; c < 3; c++) P[3 * a + c][3 * A + c] = wgt;
      }
    double T[24][24];
    for (int r = 0; r < 24; r++)
      for (int c = 0; c < 24; c++) {
        double acc = 0;
        for (int q = 0; q < 24; q++) acc += t.Ke[r * 24 + q] * P[q][c];
        T[r][c] = acc;
      }
    for (int r = 0; r < 24; r++)
      for (int c = 0; c < 24; c++) {
        double acc = 0;
        for (int q = 0; q < 24; q++) acc += P[q][r] * T[q][c];
        t.Mchild[f * 576 + r * 24 + c] = acc;
      }
  }
  return t;
}

}  // namespace femh
