/*
 * femh_oracle.c -- CPU FP64 restatement of desaic/fem's periodic voxel
 * homogenization (the checker for the HIP engine).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke() check and bench.py's
 * cpu_baseline leg load this file's library, and only as the checker or the
 * timed CPU baseline -- the product library never links or calls it.
 *
 * What each function restates (reference = desaic/fem):
 *   or_ke_hex        femlib/src/MaterialQuad.cpp:68-105 (Ke = vol*sum_q w_q K_q)
 *                    femlib/src/StrainLin.cpp:29-37 (dP = mu(dF+dF^T)+lambda tr(dF) I)
 *                    femlib/src/ElementHex.cpp:73-89 (shapeFunGrad), :24-33 (sw)
 *                    femlib/src/Quadrature.cpp:3-7,63-77 (Gauss2, w=0.125)
 *   or_apply/or_rhs  periodic assembly of coef_e*K0 over the voxel grid, strain-
 *                    driven load cases (NumericalCoarseningUsingMultiGrid.cpp:
 *                    134-288; the periodic node map of copyPeriodicU :16-41)
 *   or_pcg_*         pcgsolver/pcg_solver.h:243-285 (PCG loop; 2-norm relative
 *                    stopping test instead of the inf-norm one, Jacobi or
 *                    multigrid preconditioner instead of MIC(0))
 *   or_ch_energy     NumericalCoarseningUsingMultiGrid.cpp:290-313 and
 *                    fast_fem/FEM3DFun.cpp:159-180 (C_ij = 1/vol sum coef u_i^T K0 u_j)
 *   or_grad_E        fast_fem/FEM3DFun.cpp:222-268 (dG_ij/dx_e = coef'(x_e) u_i^T K0 u_j)
 *   or_thermal_stress SamplesGenerator/ThermalAnalysis.cpp:210-274, :388-505 (the homogenized
 *                    thermal-expansion coefficient; here by reciprocity from the
 *                    six elastic fields, see the function)
 *   multigrid        the reference's multigrid lives in topoOptLib, which is not
 *                    vendored; this is a textbook Galerkin geometric multigrid
 *                    (trilinear prolongation, R = P^T, coarse element matrices
 *                    sum_c P_c^T K_c P_c, damped Jacobi smoothing).
 */
#include "femh_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, k, ny, nz) (((size_t)(i) * (ny) + (j)) * (nz) + (k))

static inline int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

double or_wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void or_lame(double E, double nu, double *lambda, double *mu) {
  *lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  *mu = E / (2.0 * (1.0 + nu));
}

void or_ke_hex(double E, double nu, double h, int float_gauss, double *Ke) {
  static const int sw[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
                               {1, -1, -1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
  double lam, mu;
  or_lame(E, nu, &lam, &mu);
  double g = float_gauss ? (double)(1.0f / sqrtf(3.0f)) : 1.0 / sqrt(3.0);
  memset(Ke, 0, sizeof(double) * 576);
  double inv4h = 1.0 / (4.0 * h);
  for (int q = 0; q < 8; q++) {
    double xi[3] = {sw[q][0] * g, sw[q][1] * g, sw[q][2] * g};
    double gN[8][3];
    for (int a = 0; a < 8; a++) {
      gN[a][0] = sw[a][0] * inv4h * (1 + sw[a][1] * xi[1]) * (1 + sw[a][2] * xi[2]);
      gN[a][1] = sw[a][1] * inv4h * (1 + sw[a][0] * xi[0]) * (1 + sw[a][2] * xi[2]);
      gN[a][2] = sw[a][2] * inv4h * (1 + sw[a][0] * xi[0]) * (1 + sw[a][1] * xi[1]);
    }
    for (int ii = 0; ii < 8; ii++)
      for (int jj = 0; jj < 3; jj++) {
        /* dF = e_jj (x) gradN_ii ; dP = mu(dF+dF^T) + lam tr(dF) I */
        double dF[3][3] = {{0}}, dP[3][3];
        for (int m = 0; m < 3; m++) dF[jj][m] = gN[ii][m];
        double tr = dF[0][0] + dF[1][1] + dF[2][2];
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++)
            dP[r][c] = mu * (dF[r][c] + dF[c][r]) + (r == c ? lam * tr : 0.0);
        for (int vv = 0; vv < 8; vv++)
          for (int kk = 0; kk < 3; kk++) {
            double f = dP[kk][0] * gN[vv][0] + dP[kk][1] * gN[vv][1] + dP[kk][2] * gN[vv][2];
            Ke[(3 * vv + kk) * 24 + 3 * ii + jj] += 0.125 * f;
          }
      }
  }
  double vol = h * h * h;
  for (int t = 0; t < 576; t++) Ke[t] *= vol;
}

void or_macro_strain(int lc, double eps[9]) {
  memset(eps, 0, 9 * sizeof(double));
  switch (lc) {
  case 0: eps[0] = 1; break;
  case 1: eps[4] = 1; break;
  case 2: eps[8] = 1; break;
  case 3: eps[5] = eps[7] = 0.5; break; /* yz */
  case 4: eps[2] = eps[6] = 0.5; break; /* xz */
  case 5: eps[1] = eps[3] = 0.5; break; /* xy */
  }
}

/* element-local affine field of load case lc (unit voxel, corner 0 at origin) */
static void affine_local(int lc, double *u0) {
  double eps[9];
  or_macro_strain(lc, eps);
  for (int v = 0; v < 8; v++) {
    double x[3] = {(v >> 2) & 1, (v >> 1) & 1, v & 1};
    for (int c = 0; c < 3; c++)
      u0[3 * v + c] = eps[3 * c] * x[0] + eps[3 * c + 1] * x[1] + eps[3 * c + 2] * x[2];
  }
}

/* gather the 24 local dofs of element (i,j,k) from a component-major field */
static inline void gather_elem(int nx, int ny, int nz, int i, int j, int k,
                               const double *u, double *ue) {
  size_t V = (size_t)nx * ny * nz;
  for (int v = 0; v < 8; v++) {
    size_t n = IDX(wrap(i + ((v >> 2) & 1), nx), wrap(j + ((v >> 1) & 1), ny),
                   wrap(k + (v & 1), nz), ny, nz);
    ue[3 * v] = u[n];
    ue[3 * v + 1] = u[V + n];
    ue[3 * v + 2] = u[2 * V + n];
  }
}

/* Node-gather form of the periodic assembly: y(n) = sum over the 8 elements
 * touching n of coef_e * (K0 u_e)[local rows of n].  Identical to the
 * element-scatter loop of the reference, written race-free for threads. */
void or_apply(int nx, int ny, int nz, const double *E, const double *Ke,
              const double *u, double *y) {
  size_t V = (size_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < 8; a++) {
          int ei = i - ((a >> 2) & 1), ej = j - ((a >> 1) & 1), ek = k - (a & 1);
          double ue[24];
          gather_elem(nx, ny, nz, ei, ej, ek, u, ue);
          double coef = E[IDX(wrap(ei, nx), wrap(ej, ny), wrap(ek, nz), ny, nz)];
          for (int c = 0; c < 3; c++) {
            const double *row = Ke + (3 * a + c) * 24;
            double s = 0;
            for (int t = 0; t < 24; t++) s += row[t] * ue[t];
            acc[c] += coef * s;
          }
        }
        size_t n = IDX(i, j, k, ny, nz);
        y[n] = acc[0];
        y[V + n] = acc[1];
        y[2 * V + n] = acc[2];
      }
}

void or_diag(int nx, int ny, int nz, const double *E, const double *Ke, double *d) {
  size_t V = (size_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < 8; a++) {
          int ei = wrap(i - ((a >> 2) & 1), nx), ej = wrap(j - ((a >> 1) & 1), ny),
              ek = wrap(k - (a & 1), nz);
          double coef = E[IDX(ei, ej, ek, ny, nz)];
          for (int c = 0; c < 3; c++) acc[c] += coef * Ke[(3 * a + c) * 24 + 3 * a + c];
        }
        size_t n = IDX(i, j, k, ny, nz);
        d[n] = acc[0];
        d[V + n] = acc[1];
        d[2 * V + n] = acc[2];
      }
}

void or_rhs(int nx, int ny, int nz, const double *E, const double *Ke, int lc, double *b) {
  size_t V = (size_t)nx * ny * nz;
  double u0[24], f[24];
  affine_local(lc, u0);
  for (int r = 0; r < 24; r++) {
    double s = 0;
    for (int t = 0; t < 24; t++) s += Ke[r * 24 + t] * u0[t];
    f[r] = s;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < 8; a++) {
          int ei = wrap(i - ((a >> 2) & 1), nx), ej = wrap(j - ((a >> 1) & 1), ny),
              ek = wrap(k - (a & 1), nz);
          double coef = E[IDX(ei, ej, ek, ny, nz)];
          for (int c = 0; c < 3; c++) acc[c] -= coef * f[3 * a + c];
        }
        size_t n = IDX(i, j, k, ny, nz);
        b[n] = acc[0];
        b[V + n] = acc[1];
        b[2 * V + n] = acc[2];
      }
}

/* ------------------------------------------------------------------------ */
/* vector helpers                                                           */
static double dot(size_t n, const double *a, const double *b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (size_t t = 0; t < n; t++) s += a[t] * b[t];
  return s;
}

static void remove_mean(size_t V, double *x) {
  for (int c = 0; c < 3; c++) {
    double s = 0;
    for (size_t t = 0; t < V; t++) s += x[c * V + t];
    s /= (double)V;
    for (size_t t = 0; t < V; t++) x[c * V + t] -= s;
  }
}

/* ------------------------------------------------------------------------ */
/* Galerkin multigrid with element matrices on every level                  */
typedef struct {
  int L;
  int dims[16][3];
  double *K[16];    /* level >= 1: per-element 24x24 matrices */
  double *diag[16]; /* point diagonals, component-major */
  double lmax[16];  /* power-iteration estimate of lambda_max(D^-1 A) */
  const double *E;  /* level 0 coefficients */
  double Ke[576];   /* level 0 unit stiffness */
  double *dense;    /* Cholesky factor of the regularised coarsest operator */
  int ndense;
  /* scratch per level */
  double *x[16], *b[16], *r[16];
} mg_t;

static size_t lvlV(const mg_t *m, int l) {
  return (size_t)m->dims[l][0] * m->dims[l][1] * m->dims[l][2];
}

/* child-to-coarse trilinear map for child c (bits cx,cy,cz) of a coarse element */
static void child_P(int c, double *P /*24x24*/) {
  memset(P, 0, 576 * sizeof(double));
  int cb[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
  for (int a = 0; a < 8; a++) {
    int ab[3] = {(a >> 2) & 1, (a >> 1) & 1, a & 1};
    for (int A = 0; A < 8; A++) {
      int Ab[3] = {(A >> 2) & 1, (A >> 1) & 1, A & 1};
      double w = 1;
      for (int ax = 0; ax < 3; ax++) {
        double x = 0.5 * (cb[ax] + ab[ax]);
        w *= Ab[ax] ? x : 1.0 - x;
      }
      for (int comp = 0; comp < 3; comp++) P[(3 * a + comp) * 24 + 3 * A + comp] = w;
    }
  }
}

/* out += s * P^T K P  (all 24x24) */
static void ptkp_acc(const double *P, const double *K, double s, double *out) {
  double T[576];
  for (int r = 0; r < 24; r++)
    for (int c = 0; c < 24; c++) {
      double acc = 0;
      for (int t = 0; t < 24; t++) acc += K[r * 24 + t] * P[t * 24 + c];
      T[r * 24 + c] = acc;
    }
  for (int r = 0; r < 24; r++)
    for (int c = 0; c < 24; c++) {
      double acc = 0;
      for (int t = 0; t < 24; t++) acc += P[t * 24 + r] * T[t * 24 + c];
      out[r * 24 + c] += s * acc;
    }
}

static void mg_apply_level(const mg_t *m, int l, const double *u, double *y) {
  int nx = m->dims[l][0], ny = m->dims[l][1], nz = m->dims[l][2];
  if (l == 0) { or_apply(nx, ny, nz, m->E, m->Ke, u, y); return; }
  size_t V = lvlV(m, l);
  const double *Kl = m->K[l];
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < 8; a++) {
          int ei = i - ((a >> 2) & 1), ej = j - ((a >> 1) & 1), ek = k - (a & 1);
          double ue[24];
          gather_elem(nx, ny, nz, ei, ej, ek, u, ue);
          const double *Kel = Kl + 576 * IDX(wrap(ei, nx), wrap(ej, ny), wrap(ek, nz), ny, nz);
          for (int c = 0; c < 3; c++) {
            const double *row = Kel + (3 * a + c) * 24;
            double s = 0;
            for (int t = 0; t < 24; t++) s += row[t] * ue[t];
            acc[c] += s;
          }
        }
        size_t n = IDX(i, j, k, ny, nz);
        y[n] = acc[0];
        y[V + n] = acc[1];
        y[2 * V + n] = acc[2];
      }
}

void or_mg_restrict(void *mv, int l, const double *rf, double *rc) {
  mg_t *m = (mg_t *)mv;
  int fx = m->dims[l][0], fy = m->dims[l][1], fz = m->dims[l][2];
  int cx = m->dims[l + 1][0], cy = m->dims[l + 1][1], cz = m->dims[l + 1][2];
  size_t Vf = lvlV(m, l), Vc = lvlV(m, l + 1);
#pragma omp parallel for collapse(2) schedule(static)
  for (int I = 0; I < cx; I++)
    for (int J = 0; J < cy; J++)
      for (int K = 0; K < cz; K++) {
        double acc[3] = {0, 0, 0};
        for (int oi = -1; oi <= 1; oi++)
          for (int oj = -1; oj <= 1; oj++)
            for (int ok = -1; ok <= 1; ok++) {
              double w = (oi ? 0.5 : 1.0) * (oj ? 0.5 : 1.0) * (ok ? 0.5 : 1.0);
              size_t n = IDX(wrap(2 * I + oi, fx), wrap(2 * J + oj, fy), wrap(2 * K + ok, fz), fy, fz);
              for (int c = 0; c < 3; c++) acc[c] += w * rf[c * Vf + n];
            }
        size_t N = IDX(I, J, K, cy, cz);
        for (int c = 0; c < 3; c++) rc[c * Vc + N] = acc[c];
      }
}

void or_mg_prolong_add(void *mv, int l, const double *ec, double *ef) {
  mg_t *m = (mg_t *)mv;
  int fx = m->dims[l][0], fy = m->dims[l][1], fz = m->dims[l][2];
  int cx = m->dims[l + 1][0], cy = m->dims[l + 1][1], cz = m->dims[l + 1][2];
  size_t Vf = lvlV(m, l), Vc = lvlV(m, l + 1);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < fx; i++)
    for (int j = 0; j < fy; j++)
      for (int k = 0; k < fz; k++) {
        int ci[2], cj[2], ck[2];
        double wi[2], wj[2], wk[2];
        int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
        ci[0] = (i >> 1); ci[1] = wrap((i + 1) >> 1, cx); wi[0] = wi[1] = (i & 1) ? 0.5 : 1.0;
        cj[0] = (j >> 1); cj[1] = wrap((j + 1) >> 1, cy); wj[0] = wj[1] = (j & 1) ? 0.5 : 1.0;
        ck[0] = (k >> 1); ck[1] = wrap((k + 1) >> 1, cz); wk[0] = wk[1] = (k & 1) ? 0.5 : 1.0;
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < ni; a++)
          for (int b = 0; b < nj; b++)
            for (int d = 0; d < nk; d++) {
              size_t N = IDX(ci[a], cj[b], ck[d], cy, cz);
              double w = wi[a] * wj[b] * wk[d];
              for (int c = 0; c < 3; c++) acc[c] += w * ec[c * Vc + N];
            }
        size_t n = IDX(i, j, k, fy, fz);
        for (int c = 0; c < 3; c++) ef[c * Vf + n] += acc[c];
      }
}

static int chol(double *A, int n) {
  for (int j = 0; j < n; j++) {
    double s = A[(size_t)j * n + j];
    for (int k = 0; k < j; k++) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(s > 0)) return 1;
    s = sqrt(s);
    A[(size_t)j * n + j] = s;
    for (int i = j + 1; i < n; i++) {
      double t = A[(size_t)i * n + j];
      for (int k = 0; k < j; k++) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = t / s;
    }
  }
  return 0;
}

static void chol_solve(const double *L, int n, double *x) {
  for (int i = 0; i < n; i++) {
    double s = x[i];
    for (int k = 0; k < i; k++) s -= L[(size_t)i * n + k] * x[k];
    x[i] = s / L[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int k = i + 1; k < n; k++) s -= L[(size_t)k * n + i] * x[k];
    x[i] = s / L[(size_t)i * n + i];
  }
}

#define DENSE_MAX_DOF 1536
#define OR_POWER_ITERS 12

void *or_mg_create(int nx, int ny, int nz, const double *E, double nu, int levels) {
  mg_t *m = (mg_t *)calloc(1, sizeof(mg_t));
  m->E = E;
  or_ke_hex(1.0, nu, 1.0, 0, m->Ke);
  m->dims[0][0] = nx; m->dims[0][1] = ny; m->dims[0][2] = nz;
  int L = 1;
  while (L < levels && L < 16) {
    int *d = m->dims[L - 1];
    if ((d[0] & 1) || (d[1] & 1) || (d[2] & 1) || d[0] < 4 || d[1] < 4 || d[2] < 4) break;
    m->dims[L][0] = d[0] / 2; m->dims[L][1] = d[1] / 2; m->dims[L][2] = d[2] / 2;
    L++;
  }
  m->L = L;
  double Pc[8][576];
  for (int c = 0; c < 8; c++) child_P(c, Pc[c]);
  /* level 1 from fine coefficients: sum_c E_c P_c^T Ke P_c */
  double M[8][576];
  for (int c = 0; c < 8; c++) {
    memset(M[c], 0, sizeof(M[c]));
    ptkp_acc(Pc[c], m->Ke, 1.0, M[c]);
  }
  for (int l = 1; l < L; l++) {
    int cy = m->dims[l][1], cz = m->dims[l][2];
    int fy = m->dims[l - 1][1], fz = m->dims[l - 1][2];
    size_t Vc = lvlV(m, l);
    m->K[l] = (double *)calloc(Vc * 576, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long long N = 0; N < (long long)Vc; N++) {
      int I = (int)(N / ((size_t)cy * cz)), J = (int)((N / cz) % cy), K = (int)(N % cz);
      double *out = m->K[l] + 576 * (size_t)N;
      for (int c = 0; c < 8; c++) {
        size_t f = IDX(2 * I + ((c >> 2) & 1), 2 * J + ((c >> 1) & 1), 2 * K + (c & 1), fy, fz);
        if (l == 1) {
          double s = E[f];
          for (int t = 0; t < 576; t++) out[t] += s * M[c][t];
        } else {
          ptkp_acc(Pc[c], m->K[l - 1] + 576 * f, 1.0, out);
        }
      }
    }
  }
  /* diagonals */
  for (int l = 0; l < L; l++) {
    size_t V = lvlV(m, l);
    m->diag[l] = (double *)malloc(3 * V * sizeof(double));
    m->x[l] = (double *)calloc(3 * V, sizeof(double));
    m->b[l] = (double *)calloc(3 * V, sizeof(double));
    m->r[l] = (double *)calloc(3 * V, sizeof(double));
    if (l == 0) {
      or_diag(nx, ny, nz, E, m->Ke, m->diag[0]);
    } else {
      int lx = m->dims[l][0], ly = m->dims[l][1], lz = m->dims[l][2];
      for (int i = 0; i < lx; i++)
        for (int j = 0; j < ly; j++)
          for (int k = 0; k < lz; k++) {
            double acc[3] = {0, 0, 0};
            for (int a = 0; a < 8; a++) {
              size_t e = IDX(wrap(i - ((a >> 2) & 1), lx), wrap(j - ((a >> 1) & 1), ly),
                             wrap(k - (a & 1), lz), ly, lz);
              for (int c = 0; c < 3; c++) acc[c] += m->K[l][576 * e + (3 * a + c) * 24 + 3 * a + c];
            }
            size_t n = IDX(i, j, k, ly, lz);
            for (int c = 0; c < 3; c++) m->diag[l][c * V + n] = acc[c];
          }
    }
  }
  /* lambda_max(D^-1 A) per level: fixed-count power iteration from a fixed
   * start vector (deterministic), then a 10% safety margin */
  for (int l = 0; l < L; l++) {
    size_t n = 3 * lvlV(m, l);
    double *x = (double *)malloc(n * sizeof(double)), *y = (double *)malloc(n * sizeof(double));
    unsigned long long s = 0x9E3779B97F4A7C15ull;
    for (size_t t = 0; t < n; t++) {
      s ^= s << 13; s ^= s >> 7; s ^= s << 17;
      x[t] = (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    double lam = 0;
    for (int it = 0; it < OR_POWER_ITERS; it++) {
      double xn = sqrt(dot(n, x, x));
      for (size_t t = 0; t < n; t++) x[t] /= xn;
      mg_apply_level(m, l, x, y);
      for (size_t t = 0; t < n; t++) y[t] /= m->diag[l][t];
      lam = sqrt(dot(n, y, y));
      double *tmp = x; x = y; y = tmp;
    }
    m->lmax[l] = 1.1 * lam;
    free(x); free(y);
  }
  /* dense coarsest factor when small enough */
  size_t Vc = lvlV(m, L - 1);
  if (3 * Vc <= DENSE_MAX_DOF && L > 1) {
    int n = (int)(3 * Vc);
    m->ndense = n;
    m->dense = (double *)calloc((size_t)n * n, sizeof(double));
    double *e = (double *)calloc(n, sizeof(double));
    double *col = (double *)calloc(n, sizeof(double));
    double dmean = 0;
    for (int t = 0; t < n; t++) dmean += m->diag[L - 1][t];
    dmean /= n;
    for (int t = 0; t < n; t++) {
      memset(e, 0, n * sizeof(double));
      e[t] = 1;
      mg_apply_level(m, L - 1, e, col);
      for (int s = 0; s < n; s++) m->dense[(size_t)s * n + t] = col[s];
    }
    /* regularise the translation null space: + dmean/Vc * z_c z_c^T */
    for (int c = 0; c < 3; c++)
      for (size_t a = 0; a < Vc; a++)
        for (size_t b = 0; b < Vc; b++)
          m->dense[(c * Vc + a) * n + c * Vc + b] += dmean / (double)Vc;
    if (chol(m->dense, n)) { free(m->dense); m->dense = NULL; m->ndense = 0; }
    free(e);
    free(col);
  }
  return m;
}

void or_mg_destroy(void *mv) {
  mg_t *m = (mg_t *)mv;
  if (!m) return;
  for (int l = 0; l < m->L; l++) {
    free(m->K[l]); free(m->diag[l]); free(m->x[l]); free(m->b[l]); free(m->r[l]);
  }
  free(m->dense);
  free(m);
}

int or_mg_level_dims(void *mv, int l, int *dims) {
  mg_t *m = (mg_t *)mv;
  if (l < 0 || l >= m->L) return m->L;
  dims[0] = m->dims[l][0]; dims[1] = m->dims[l][1]; dims[2] = m->dims[l][2];
  return m->L;
}

void or_mg_apply(void *mv, int l, const double *u, double *y) {
  mg_apply_level((mg_t *)mv, l, u, y);
}

static void jacobi_pcg_level(const mg_t *m, int l, const double *b, double *x, int iters) {
  size_t n = 3 * lvlV(m, l);
  double *r = (double *)malloc(n * sizeof(double)), *z = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * sizeof(double)), *q = (double *)malloc(n * sizeof(double));
  memset(x, 0, n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  for (size_t t = 0; t < n; t++) p[t] = z[t] = r[t] / m->diag[l][t];
  double rz = dot(n, r, z);
  for (int it = 0; it < iters && rz > 0; it++) {
    mg_apply_level(m, l, p, q);
    double pq = dot(n, p, q);
    if (!(pq > 0)) break;
    double al = rz / pq;
    for (size_t t = 0; t < n; t++) { x[t] += al * p[t]; r[t] -= al * q[t]; }
    for (size_t t = 0; t < n; t++) z[t] = r[t] / m->diag[l][t];
    double rzn = dot(n, r, z);
    double be = rzn / rz;
    rz = rzn;
    for (size_t t = 0; t < n; t++) p[t] = z[t] + be * p[t];
  }
  free(r); free(z); free(p); free(q);
}

/* Smoother: x <- x + p(D^-1 A) D^-1 (b - A x), either damped Jacobi with
 * omega/lambda_max per sweep, or a Chebyshev polynomial of degree `sweeps`
 * on [lambda_max/cheby_ratio, lambda_max].  x_is_zero skips the first
 * residual apply. */
static void smooth(mg_t *m, int l, const double *b, double *x, int sweeps, int x_is_zero,
                   const or_solver_opts *o) {
  size_t n = 3 * lvlV(m, l);
  const double *D = m->diag[l];
  double *r = m->r[l];
  double *d = (double *)malloc(n * sizeof(double));
  if (x_is_zero) { memcpy(r, b, n * sizeof(double)); memset(x, 0, n * sizeof(double)); }
  else { mg_apply_level(m, l, x, r); for (size_t t = 0; t < n; t++) r[t] = b[t] - r[t]; }
  double lmax = m->lmax[l];
  if (o->smoother == 0) {
    double w = o->omega / lmax;
    for (int s = 0; s < sweeps; s++) {
      for (size_t t = 0; t < n; t++) x[t] += w * r[t] / D[t];
      if (s + 1 < sweeps) { mg_apply_level(m, l, x, r); for (size_t t = 0; t < n; t++) r[t] = b[t] - r[t]; }
    }
  } else {
    double lo = lmax / o->cheby_ratio, hi = lmax;
    double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    for (size_t t = 0; t < n; t++) d[t] = r[t] / (D[t] * theta);
    double *Ad = (double *)malloc(n * sizeof(double));
    for (int s = 0; s < sweeps; s++) {
      for (size_t t = 0; t < n; t++) x[t] += d[t];
      if (s + 1 == sweeps) break;
      mg_apply_level(m, l, d, Ad);
      for (size_t t = 0; t < n; t++) r[t] -= Ad[t];
      double rho_new = 1.0 / (2.0 * sigma - rho);
      for (size_t t = 0; t < n; t++) d[t] = rho_new * rho * d[t] + (2.0 * rho_new / delta) * r[t] / D[t];
      rho = rho_new;
    }
    free(Ad);
  }
  free(d);
}

static void vcycle(mg_t *m, int l, const double *b, double *x, const or_solver_opts *o) {
  size_t n = 3 * lvlV(m, l);
  if (l == m->L - 1) {
    if (m->dense) {
      memcpy(x, b, n * sizeof(double));
      chol_solve(m->dense, m->ndense, x);
    } else {
      jacobi_pcg_level(m, l, b, x, o->coarse_iters);
    }
    return;
  }
  double *r = m->r[l];
  smooth(m, l, b, x, o->nu1, 1, o);
  mg_apply_level(m, l, x, r);
  for (size_t t = 0; t < n; t++) r[t] = b[t] - r[t];
  or_mg_restrict(m, l, r, m->b[l + 1]);
  vcycle(m, l + 1, m->b[l + 1], m->x[l + 1], o);
  or_mg_prolong_add(m, l, m->x[l + 1], x);
  smooth(m, l, b, x, o->nu2, 0, o);
}

void or_default_opts(or_solver_opts *o) {
  o->precond = 1;
  o->levels = 16;
  o->nu1 = 2;
  o->nu2 = 2;
  o->omega = 1.0;
  o->smoother = 1;
  o->cheby_ratio = 8.0;
  o->coarse_iters = 64;
  o->tol = 1e-10;
  o->max_iter = 5000;
}

/* PCG (flexible Polak-Ribiere beta when the preconditioner is not a fixed
 * linear operator, i.e. multigrid with an iterative coarsest solve). */
static int pcg(int nx, int ny, int nz, const double *E, const double *Ke, mg_t *mg,
               const double *diag, const or_solver_opts *o, const double *b, double *x,
               int *iters_out, double *relres_out) {
  size_t V = (size_t)nx * ny * nz, n = 3 * V;
  double *r = (double *)malloc(n * sizeof(double)), *z = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * sizeof(double)), *q = (double *)malloc(n * sizeof(double));
  double *zold = NULL;
  int flexible = (o->precond == 1 && mg && !mg->dense && mg->L > 1);
  if (flexible) zold = (double *)calloc(n, sizeof(double));
  memset(x, 0, n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  double bn = sqrt(dot(n, b, b));
  int it = 0, conv = 0;
  double rn = bn;
  if (bn == 0) { conv = 1; goto done; }
#define PRECOND(src, dst)                                                   \
  do {                                                                      \
    if (o->precond == 1 && mg) vcycle(mg, 0, src, dst, o);                  \
    else for (size_t t = 0; t < n; t++) dst[t] = src[t] / diag[t];          \
  } while (0)
  PRECOND(r, z);
  memcpy(p, z, n * sizeof(double));
  double rz = dot(n, r, z);
  for (it = 0; it < o->max_iter;) {
    or_apply(nx, ny, nz, E, Ke, p, q);
    double pq = dot(n, p, q);
    double al = rz / pq;
    for (size_t t = 0; t < n; t++) { x[t] += al * p[t]; r[t] -= al * q[t]; }
    it++;
    rn = sqrt(dot(n, r, r));
    if (rn <= o->tol * bn) { conv = 1; break; }
    if (flexible) memcpy(zold, z, n * sizeof(double));
    PRECOND(r, z);
    double rzn = dot(n, r, z);
    double be;
    if (flexible) {
      double s = 0;
      for (size_t t = 0; t < n; t++) s += r[t] * zold[t];
      be = (rzn - s) / rz;
    } else {
      be = rzn / rz;
    }
    rz = rzn;
    for (size_t t = 0; t < n; t++) p[t] = z[t] + be * p[t];
  }
#undef PRECOND
done:
  remove_mean(V, x);
  *iters_out = it;
  *relres_out = bn > 0 ? rn / bn : 0.0;
  free(r); free(z); free(p); free(q); free(zold);
  return conv ? 0 : 1;
}

/* Effective thermal stress per unit temperature rise: beta_q = (1/V) sum_e
 * alpha_e E_e int_e I : C0 : eps(chi_q), chi_q = affine + fluctuation of load
 * case q.  With the elastic fields this is the exact periodic thermal stress
 * of the discretisation (reciprocity), so no thermal solve is needed; the
 * reference estimates the same quantity by a force-driven solve per cell
 * (SamplesGenerator/ThermalAnalysis.cpp:388-505, computeThermalCoefficient)
 * and by Levin's formula (:263-266).  The strain is evaluated at the 8 Gauss
 * points of or_ke_hex and the stress by the Lame form. */
void or_thermal_stress(int nx, int ny, int nz, const double *E, double nu, const double *alpha,
                       const double *u, double *beta) {
  static const int sw[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
                               {1, -1, -1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
  size_t V = (size_t)nx * ny * nz;
  double lam, mu;
  or_lame(1.0, nu, &lam, &mu);
  const double g = 1.0 / sqrt(3.0);
  double gN[8][8][3]; /* [gauss point][corner][axis] */
  for (int q = 0; q < 8; q++)
    for (int a = 0; a < 8; a++) {
      double xi[3] = {sw[q][0] * g, sw[q][1] * g, sw[q][2] * g};
      gN[q][a][0] = 0.25 * sw[a][0] * (1 + sw[a][1] * xi[1]) * (1 + sw[a][2] * xi[2]);
      gN[q][a][1] = 0.25 * sw[a][1] * (1 + sw[a][0] * xi[0]) * (1 + sw[a][2] * xi[2]);
      gN[q][a][2] = 0.25 * sw[a][2] * (1 + sw[a][0] * xi[0]) * (1 + sw[a][1] * xi[1]);
    }
  double u0[6][24];
  for (int lc = 0; lc < 6; lc++) affine_local(lc, u0[lc]);
  double acc[6] = {0};
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        size_t e = IDX(i, j, k, ny, nz);
        double w = alpha[e] * E[e];
        if (w == 0.0) continue;
        for (int lc = 0; lc < 6; lc++) {
          double chi[24];
          gather_elem(nx, ny, nz, i, j, k, u + 3 * V * lc, chi);
          for (int t = 0; t < 24; t++) chi[t] += u0[lc][t];
          double s = 0;
          for (int q = 0; q < 8; q++) {
            double F[3][3] = {{0}};
            for (int a = 0; a < 8; a++)
              for (int c = 0; c < 3; c++)
                for (int m = 0; m < 3; m++) F[c][m] += chi[3 * a + c] * gN[q][a][m];
            double tr = F[0][0] + F[1][1] + F[2][2], trs = 0;
            for (int c = 0; c < 3; c++) trs += lam * tr + 2.0 * mu * F[c][c]; /* I : sigma */
            s += 0.125 * trs;
          }
          acc[lc] += w * s;
        }
      }
  for (int lc = 0; lc < 6; lc++) beta[lc] = acc[lc] / (double)V;
}

void or_ch_energy(int nx, int ny, int nz, const double *E, double nu, const double *u, double *C) {
  double Ke[576];
  or_ke_hex(1.0, nu, 1.0, 0, Ke);
  size_t V = (size_t)nx * ny * nz;
  double u0[6][24];
  for (int lc = 0; lc < 6; lc++) affine_local(lc, u0[lc]);
  double acc[36] = {0};
#pragma omp parallel
  {
    double loc[36] = {0};
#pragma omp for collapse(2) schedule(static)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          double chi[6][24], Kc[6][24];
          for (int lc = 0; lc < 6; lc++) {
            gather_elem(nx, ny, nz, i, j, k, u + 3 * V * lc, chi[lc]);
            for (int t = 0; t < 24; t++) chi[lc][t] += u0[lc][t];
            for (int r = 0; r < 24; r++) {
              double s = 0;
              for (int t = 0; t < 24; t++) s += Ke[r * 24 + t] * chi[lc][t];
              Kc[lc][r] = s;
            }
          }
          double coef = E[IDX(i, j, k, ny, nz)];
          for (int a = 0; a < 6; a++)
            for (int b = a; b < 6; b++) {
              double s = 0;
              for (int t = 0; t < 24; t++) s += chi[a][t] * Kc[b][t];
              loc[6 * a + b] += coef * s;
            }
        }
#pragma omp critical
    for (int t = 0; t < 36; t++) acc[t] += loc[t];
  }
  for (int a = 0; a < 6; a++)
    for (int b = a; b < 6; b++) C[6 * a + b] = C[6 * b + a] = acc[6 * a + b] / (double)V;
}

void or_energy_terms(int nx, int ny, int nz, double nu, const double *u, double *W) {
  double Ke[576];
  or_ke_hex(1.0, nu, 1.0, 0, Ke);
  size_t V = (size_t)nx * ny * nz;
  double u0[6][24];
  for (int lc = 0; lc < 6; lc++) affine_local(lc, u0[lc]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double chi[6][24], Kc[6][24];
        for (int lc = 0; lc < 6; lc++) {
          gather_elem(nx, ny, nz, i, j, k, u + 3 * V * lc, chi[lc]);
          for (int t = 0; t < 24; t++) chi[lc][t] += u0[lc][t];
          for (int r = 0; r < 24; r++) {
            double s = 0;
            for (int t = 0; t < 24; t++) s += Ke[r * 24 + t] * chi[lc][t];
            Kc[lc][r] = s;
          }
        }
        size_t e = IDX(i, j, k, ny, nz);
        int p = 0;
        for (int a = 0; a < 6; a++)
          for (int b = a; b < 6; b++) {
            double s = 0;
            for (int t = 0; t < 24; t++) s += chi[a][t] * Kc[b][t];
            W[21 * e + p++] = s / (double)V;
          }
      }
}

void or_grad_E(int nx, int ny, int nz, double nu, const double *u, const double *G, double *dJdE) {
  size_t V = (size_t)nx * ny * nz;
  double *W = (double *)malloc(21 * V * sizeof(double));
  or_energy_terms(nx, ny, nz, nu, u, W);
  double Gs[21];
  int p = 0;
  for (int a = 0; a < 6; a++)
    for (int b = a; b < 6; b++) Gs[p++] = (a == b) ? G[6 * a + a] : G[6 * a + b] + G[6 * b + a];
  for (size_t e = 0; e < V; e++) {
    double s = 0;
    for (int t = 0; t < 21; t++) s += Gs[t] * W[21 * e + t];
    dJdE[e] = s;
  }
  free(W);
}

int or_homogenize_cases(int nx, int ny, int nz, const double *E, double nu, const or_solver_opts *o,
                        int case_mask, double *C, double *u, int *iters, double *relres,
                        double *seconds) {
  const double t0 = or_wall();
  size_t V = (size_t)nx * ny * nz;
  double Ke[576];
  or_ke_hex(1.0, nu, 1.0, 0, Ke);
  double *diag = (double *)malloc(3 * V * sizeof(double));
  or_diag(nx, ny, nz, E, Ke, diag);
  mg_t *mg = NULL;
  if (o->precond == 1) mg = (mg_t *)or_mg_create(nx, ny, nz, E, nu, o->levels);
  double *b = (double *)malloc(3 * V * sizeof(double));
  /* scale below which a right-hand side counts as exactly zero (homogeneous
   * cells: the affine field already satisfies equilibrium) */
  double emax = 0;
  for (size_t t = 0; t < V; t++) emax = E[t] > emax ? E[t] : emax;
  const double t1 = or_wall();
  int status = 0;
  for (int lc = 0; lc < 6; lc++) {
    double *x = u + 3 * V * lc;
    if (!((case_mask >> lc) & 1)) {
      memset(x, 0, 3 * V * sizeof(double));
      iters[lc] = 0;
      relres[lc] = 0;
      continue;
    }
    or_rhs(nx, ny, nz, E, Ke, lc, b);
    double bn = sqrt(dot(3 * V, b, b));
    if (bn <= 1e-13 * emax * sqrt((double)V)) {
      memset(x, 0, 3 * V * sizeof(double));
      iters[lc] = 0;
      relres[lc] = 0;
      continue;
    }
    status |= pcg(nx, ny, nz, E, Ke, mg, diag, o, b, x, &iters[lc], &relres[lc]);
  }
  const double t2 = or_wall();
  if ((case_mask & 63) == 63) or_ch_energy(nx, ny, nz, E, nu, u, C);
  or_mg_destroy(mg);
  free(diag);
  free(b);
  if (seconds) {
    seconds[0] = t1 - t0;               /* set-up: diagonal, multigrid hierarchy */
    seconds[1] = t2 - t1;               /* the masked load-case solves */
    seconds[2] = or_wall() - t2;        /* C_H */
  }
  return status;
}

int or_homogenize(int nx, int ny, int nz, const double *E, double nu, const or_solver_opts *o,
                  double *C, double *u, int *iters, double *relres) {
  return or_homogenize_cases(nx, ny, nz, E, nu, o, 63, C, u, iters, relres, NULL);
}

/* ======================================================================
 * Force-driven periodic homogenization (the reference's NumericalCoarsening
 * path) on the (n+1)^3-vertex mesh, restated for small grids:
 *   stiffness   element K = E_e Ke(h = 1/max n) assembled on the vertices
 *               (SamplesGenerator/NumericalCoarsening.cpp:1238-1306); with
 *               `regularize` every diagonal entry is scaled by 1 + 1e-5
 *               (:1277-1279)
 *   constraints translation (3 rows), rotation (3 rows) and periodicity
 *               u(v+) - u(v-) = u(c_a) - u(c_0) per face-vertex pair and
 *               axis, the pairs already tied by a lower axis left out
 *               (femlib/src/EMeshStiffness.cpp:53-274)
 *   loads       the 6 face-traction cases of computeHarmonicDisplacements
 *               (NumericalCoarsening.cpp:1152-1160, getExternalForces
 *               :941-989): per element face vertex a force 1/(n_b n_c)/4
 *               times the magnitude, - on the low and + on the high face;
 *               order xx, yy, zz, xy, yz, xz
 *   solve       the symmetric KKT system, dense, by LU with partial pivoting
 *               (the reference factorizes it with PARDISO)
 *   tensor      the coarse-strain form of femlib/src/MeshUtilities.cpp:960-
 *               1060: G = tensor strains of the six solutions at the centre
 *               of the cell-sized hex, Gv = tensor strains at the 8 Gauss
 *               points of every element, C = (1/nelem) G^-T (sum_e sum_q w_q
 *               Gv^T C_e Gv) G^-1 with C_e the engineering-shear stiffness
 *               (femlib/src/StrainLin.cpp:39-58), Voigt (xx,yy,zz,yz,xz,xy)
 * ====================================================================== */

static int vidx(int ny, int nz, int i, int j, int k) { return (i * (ny + 1) + j) * (nz + 1) + k; }

/* 6x6 engineering-shear isotropic stiffness of modulus E */
static void emat(double E, double nu, double C[36]) {
  double lam, mu;
  or_lame(E, nu, &lam, &mu);
  memset(C, 0, 36 * sizeof(double));
  for (int a = 0; a < 3; a++) {
    for (int b = 0; b < 3; b++) C[6 * a + b] = lam;
    C[6 * a + a] = lam + 2 * mu;
    C[6 * (a + 3) + a + 3] = mu;
  }
}

/* in-place Gauss-Jordan inverse of an n x n matrix; returns 0 on success */
static int invert_small(int n, double *A) {
  double *M = (double *)malloc(sizeof(double) * n * 2 * n);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < 2 * n; c++) M[r * 2 * n + c] = c < n ? A[r * n + c] : (c - n == r ? 1.0 : 0.0);
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(M[r * 2 * n + c]) > fabs(M[p * 2 * n + c])) p = r;
    if (M[p * 2 * n + c] == 0.0) {
      free(M);
      return 1;
    }
    if (p != c)
      for (int t = 0; t < 2 * n; t++) {
        double s = M[c * 2 * n + t];
        M[c * 2 * n + t] = M[p * 2 * n + t];
        M[p * 2 * n + t] = s;
      }
    double d = M[c * 2 * n + c];
    for (int t = 0; t < 2 * n; t++) M[c * 2 * n + t] /= d;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      double f = M[r * 2 * n + c];
      if (f != 0.0)
        for (int t = 0; t < 2 * n; t++) M[r * 2 * n + t] -= f * M[c * 2 * n + t];
    }
  }
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) A[r * n + c] = M[r * 2 * n + n + c];
  free(M);
  return 0;
}

/* dense LU with partial pivoting of the n x n matrix A (in place), then the
 * solution of A x = b for nrhs right-hand sides (columns of b, n x nrhs
 * row-major, overwritten by x).  Returns 0 on success. */
static int lu_solve(int n, double *A, int nrhs, double *b) {
  int *piv = (int *)malloc(sizeof(int) * n);
  for (int c = 0; c < n; c++) {
    int p = c;
    double best = fabs(A[(size_t)c * n + c]);
    for (int r = c + 1; r < n; r++)
      if (fabs(A[(size_t)r * n + c]) > best) {
        best = fabs(A[(size_t)r * n + c]);
        p = r;
      }
    piv[c] = p;
    if (best == 0.0) {
      free(piv);
      return 1;
    }
    if (p != c) {
      for (int t = 0; t < n; t++) {
        double s = A[(size_t)c * n + t];
        A[(size_t)c * n + t] = A[(size_t)p * n + t];
        A[(size_t)p * n + t] = s;
      }
      for (int t = 0; t < nrhs; t++) {
        double s = b[(size_t)c * nrhs + t];
        b[(size_t)c * nrhs + t] = b[(size_t)p * nrhs + t];
        b[(size_t)p * nrhs + t] = s;
      }
    }
    const double d = A[(size_t)c * n + c];
#pragma omp parallel for schedule(static)
    for (int r = c + 1; r < n; r++) {
      const double f = A[(size_t)r * n + c] / d;
      A[(size_t)r * n + c] = f;
      if (f == 0.0) continue;
      for (int t = c + 1; t < n; t++) A[(size_t)r * n + t] -= f * A[(size_t)c * n + t];
      for (int t = 0; t < nrhs; t++) b[(size_t)r * nrhs + t] -= f * b[(size_t)c * nrhs + t];
    }
  }
  for (int r = n - 1; r >= 0; r--)
    for (int t = 0; t < nrhs; t++) {
      double s = b[(size_t)r * nrhs + t];
      for (int c = r + 1; c < n; c++) s -= A[(size_t)r * n + c] * b[(size_t)c * nrhs + t];
      b[(size_t)r * nrhs + t] = s / A[(size_t)r * n + r];
    }
  free(piv);
  return 0;
}

/* vertex spacing 1/max n; femlib computes it in float (ElementRegGrid.cpp:116) */
static double grid_h(int nx, int ny, int nz, int faithful) {
  const int m = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
  return faithful ? (double)(1.0f / (float)m) : 1.0 / (double)m;
}

int or_kkt_constraints(int nx, int ny, int nz, int faithful, double *rows, int max_rows) {
  const int NV = (nx + 1) * (ny + 1) * (nz + 1), ND = 3 * NV;
  const double h = grid_h(nx, ny, nz, faithful);
  int nr = 0;
#define ROW(r) (rows + (size_t)(r) * ND)
  int need = 6 + 3 * ((ny + 1) * (nz + 1) - 1 + nx * (nz + 1) - 1 + nx * ny - 1);
  if (!rows) return need;
  if (max_rows < need) return -need;
  memset(rows, 0, sizeof(double) * (size_t)need * ND);
  /* translation: -sum_v u_v = 0 */
  for (int kk = 0; kk < 3; kk++, nr++)
    for (int v = 0; v < NV; v++) ROW(nr)[3 * v + kk] = -1.0;
  /* rotation: sum_v c_v u_v = 0, c_v = -skew(X_v - centre) */
  double cen[3] = {0, 0, 0};
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++) {
        cen[0] += i * h;
        cen[1] += j * h;
        cen[2] += k * h;
      }
  for (int a = 0; a < 3; a++) cen[a] /= NV;
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++) {
        const int v = vidx(ny, nz, i, j, k);
        const double r0 = i * h - cen[0], r1 = j * h - cen[1], r2 = k * h - cen[2];
        const double c[3][3] = {{0, r2, -r1}, {-r2, 0, r0}, {r1, -r0, 0}};
        for (int kk = 0; kk < 3; kk++)
          for (int ll = 0; ll < 3; ll++) ROW(nr + kk)[3 * v + ll] = c[kk][ll];
      }
  nr += 3;
  /* periodicity: -u(c0) + u(c_a) + u(v-) - u(v+) = 0 per component */
  const int c0 = vidx(ny, nz, 0, 0, 0);
  const int ca[3] = {vidx(ny, nz, nx, 0, 0), vidx(ny, nz, 0, ny, 0), vidx(ny, nz, 0, 0, nz)};
  for (int ax = 0; ax < 3; ax++) {
    const int na = ax == 0 ? ny + 1 : nx, nb = ax == 2 ? ny : nz + 1;
    for (int p = 0; p < na; p++)
      for (int q = 0; q < nb; q++) {
        if (p == 0 && q == 0) continue;
        int vm, vp;
        if (ax == 0) {
          vm = vidx(ny, nz, 0, p, q);
          vp = vidx(ny, nz, nx, p, q);
        } else if (ax == 1) {
          vm = vidx(ny, nz, p, 0, q);
          vp = vidx(ny, nz, p, ny, q);
        } else {
          vm = vidx(ny, nz, p, q, 0);
          vp = vidx(ny, nz, p, q, nz);
        }
        for (int kk = 0; kk < 3; kk++, nr++) {
          ROW(nr)[3 * c0 + kk] += -1.0;
          ROW(nr)[3 * ca[ax] + kk] += 1.0;
          ROW(nr)[3 * vm + kk] += 1.0;
          ROW(nr)[3 * vp + kk] += -1.0;
        }
      }
  }
#undef ROW
  return nr;
}

/* coarse-strain tensor from the six vertex displacement fields u[6][3 NV]
 * (MeshUtilities.cpp:960-1060) */
void or_coarse_strain_tensor(int nx, int ny, int nz, const double *E, double nu, int faithful,
                             const double *u, double *C) {
  const int NV = (nx + 1) * (ny + 1) * (nz + 1);
  const double h = grid_h(nx, ny, nz, faithful);
  const int float_gauss = faithful;
  const double L[3] = {nx * h, ny * h, nz * h};
  /* G: cell-sized hex at its centre; columns = load cases */
  double G[36] = {0};
  for (int d = 0; d < 6; d++) {
    double gu[3][3] = {{0}};
    for (int v = 0; v < 8; v++) {
      const int b[3] = {(v >> 2) & 1, (v >> 1) & 1, v & 1};
      const int vx = vidx(ny, nz, b[0] * nx, b[1] * ny, b[2] * nz);
      double gN[3];
      for (int a = 0; a < 3; a++) gN[a] = (b[a] ? 1.0 : -1.0) / (4.0 * L[a]);
      for (int c = 0; c < 3; c++)
        for (int a = 0; a < 3; a++) gu[c][a] += u[(size_t)d * 3 * NV + 3 * vx + c] * gN[a];
    }
    G[0 * 6 + d] = gu[0][0];
    G[1 * 6 + d] = gu[1][1];
    G[2 * 6 + d] = gu[2][2];
    G[3 * 6 + d] = 0.5 * (gu[1][2] + gu[2][1]);
    G[4 * 6 + d] = 0.5 * (gu[0][2] + gu[2][0]);
    G[5 * 6 + d] = 0.5 * (gu[0][1] + gu[1][0]);
  }
  const double g = float_gauss ? (double)(1.0f / sqrtf(3.0f)) : 1.0 / sqrt(3.0);
  double S[36] = {0};
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double Ce[36];
        emat(E[(i * ny + j) * nz + k], nu, Ce);
        for (int q = 0; q < 8; q++) {
          const double xi[3] = {(q & 4) ? g : -g, (q & 2) ? g : -g, (q & 1) ? g : -g};
          double gN[8][3];
          for (int v = 0; v < 8; v++) {
            const int sgn[3] = {((v >> 2) & 1) ? 1 : -1, ((v >> 1) & 1) ? 1 : -1, (v & 1) ? 1 : -1};
            gN[v][0] = sgn[0] / (4.0 * h) * (1 + sgn[1] * xi[1]) * (1 + sgn[2] * xi[2]);
            gN[v][1] = sgn[1] / (4.0 * h) * (1 + sgn[0] * xi[0]) * (1 + sgn[2] * xi[2]);
            gN[v][2] = sgn[2] / (4.0 * h) * (1 + sgn[0] * xi[0]) * (1 + sgn[1] * xi[1]);
          }
          double Gv[36];
          for (int d = 0; d < 6; d++) {
            double gu[3][3] = {{0}};
            for (int v = 0; v < 8; v++) {
              const int vx = vidx(ny, nz, i + ((v >> 2) & 1), j + ((v >> 1) & 1), k + (v & 1));
              for (int c = 0; c < 3; c++)
                for (int a = 0; a < 3; a++) gu[c][a] += u[(size_t)d * 3 * NV + 3 * vx + c] * gN[v][a];
            }
            Gv[0 * 6 + d] = gu[0][0];
            Gv[1 * 6 + d] = gu[1][1];
            Gv[2 * 6 + d] = gu[2][2];
            Gv[3 * 6 + d] = 0.5 * (gu[1][2] + gu[2][1]);
            Gv[4 * 6 + d] = 0.5 * (gu[0][2] + gu[2][0]);
            Gv[5 * 6 + d] = 0.5 * (gu[0][1] + gu[1][0]);
          }
          for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) {
              double s = 0;
              for (int r = 0; r < 6; r++)
                for (int t = 0; t < 6; t++) s += Gv[r * 6 + a] * Ce[r * 6 + t] * Gv[t * 6 + b];
              S[a * 6 + b] += 0.125 * s;
            }
        }
      }
  double Gi[36];
  memcpy(Gi, G, sizeof(Gi));
  if (invert_small(6, Gi)) {
    for (int t = 0; t < 36; t++) C[t] = NAN;
    return;
  }
  /* femlib scales by (1.f / nelem) in float (MeshUtilities.cpp:1059) */
  const double inel = faithful ? (double)(1.0f / (float)(nx * ny * nz)) : 1.0 / ((double)nx * ny * nz);
  for (int a = 0; a < 6; a++)
    for (int b = 0; b < 6; b++) {
      double s = 0;
      for (int r = 0; r < 6; r++)
        for (int t = 0; t < 6; t++) s += Gi[r * 6 + a] * S[r * 6 + t] * Gi[t * 6 + b];
      C[a * 6 + b] = s * inel;
    }
}

int or_force_driven(int nx, int ny, int nz, const double *E, double nu, int regularize, int faithful,
                    double *C, double *u) {
  const int NV = (nx + 1) * (ny + 1) * (nz + 1), ND = 3 * NV;
  const double h = grid_h(nx, ny, nz, faithful);
  const int ncon = or_kkt_constraints(nx, ny, nz, faithful, NULL, 0);
  const int n = ND + ncon;
  double *A = (double *)calloc((size_t)n * n, sizeof(double));
  double *rows = (double *)malloc(sizeof(double) * (size_t)ncon * ND);
  double *b = (double *)calloc((size_t)n * 6, sizeof(double));
  if (!A || !rows || !b) {
    free(A);
    free(rows);
    free(b);
    return 2;
  }
  double Ke[576];
  or_ke_hex(1.0, nu, h, faithful, Ke);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const double Ee = E[(i * ny + j) * nz + k];
        int vv[8];
        for (int v = 0; v < 8; v++) vv[v] = vidx(ny, nz, i + ((v >> 2) & 1), j + ((v >> 1) & 1), k + (v & 1));
        for (int a = 0; a < 8; a++)
          for (int ca = 0; ca < 3; ca++)
            for (int bb = 0; bb < 8; bb++)
              for (int cb = 0; cb < 3; cb++)
                A[(size_t)(3 * vv[a] + ca) * n + 3 * vv[bb] + cb] += Ee * Ke[(3 * a + ca) * 24 + 3 * bb + cb];
      }
  if (regularize)
    for (int t = 0; t < ND; t++) A[(size_t)t * n + t] *= 1.0 + 1e-5;
  or_kkt_constraints(nx, ny, nz, faithful, rows, ncon);
  for (int r = 0; r < ncon; r++)
    for (int c = 0; c < ND; c++) {
      const double v = rows[(size_t)r * ND + c];
      A[(size_t)(ND + r) * n + c] = v;
      A[(size_t)c * n + ND + r] = v;
    }
  /* loads: (fx, fy, fz) per case, reference order xx yy zz xy yz xz; a
   * force along direction `dir` on the faces normal to `ax` */
  static const int pairs[6][2][2] = {{{0, 0}, {-1, -1}}, {{1, 1}, {-1, -1}}, {{2, 2}, {-1, -1}},
                                     {{1, 0}, {0, 1}},   {{2, 1}, {1, 2}},   {{2, 0}, {0, 2}}};
  const int nn[3] = {nx, ny, nz};
  for (int d = 0; d < 6; d++)
    for (int t = 0; t < 2; t++) {
      const int ax = pairs[d][t][0], dir = pairs[d][t][1];
      if (ax < 0) continue;
      const double mag = (d < 3) ? 1.0 : 0.5;
      const double coeff = mag / ((double)nn[(ax + 1) % 3] * nn[(ax + 2) % 3]) / 4.0;
      for (int side = 0; side < 2; side++) {
        const double sign = side ? 1.0 : -1.0;
        for (int i = 0; i < nx; i++)
          for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
              const int ijk[3] = {i, j, k};
              if (ijk[ax] != (side ? nn[ax] - 1 : 0)) continue;
              for (int v = 0; v < 8; v++) {
                const int bv[3] = {(v >> 2) & 1, (v >> 1) & 1, v & 1};
                if (bv[ax] != side) continue;
                const int vx = vidx(ny, nz, i + bv[0], j + bv[1], k + bv[2]);
                b[(size_t)(3 * vx + dir) * 6 + d] += sign * coeff;
              }
            }
      }
    }
  const int st = lu_solve(n, A, 6, b);
  if (!st) {
    for (int d = 0; d < 6; d++)
      for (int t = 0; t < ND; t++) u[(size_t)d * ND + t] = b[(size_t)t * 6 + d];
    or_coarse_strain_tensor(nx, ny, nz, E, nu, faithful, u, C);
  }
  free(A);
  free(rows);
  free(b);
  return st;
}

/* The same coarse-strain tensor from the strain-driven periodic fields
 * u[6][3][V] (zero-mean fluctuations of the unit engineering strains, as
 * or_homogenize returns them): with a = the engineering strains of the six
 * total fields at a Gauss point, T = diag(1,1,1,1/2,1/2,1/2) (engineering to
 * tensor shear) and C0 the unit-modulus engineering stiffness,
 * C = T^-1 [ (1/V) sum_e E_e sum_q w_q a^T T C0 T a ] T^-1.  This is what the
 * force-driven tensor reduces to when the constrained solutions are affine
 * plus periodic (no diagonal regularisation). */
void or_coarse_strain_from_fields(int nx, int ny, int nz, const double *E, double nu, int float_gauss,
                                  const double *u, double *C) {
  const size_t V = (size_t)nx * ny * nz;
  double C0[36], D[36];
  emat(1.0, nu, C0);
  const double Tv[6] = {1, 1, 1, 0.5, 0.5, 0.5};
  for (int a = 0; a < 6; a++)
    for (int b = 0; b < 6; b++) D[a * 6 + b] = Tv[a] * C0[a * 6 + b] * Tv[b];
  const double g = float_gauss ? (double)(1.0f / sqrtf(3.0f)) : 1.0 / sqrt(3.0);
  double M[36] = {0};
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const double Ee = E[(i * ny + j) * nz + k];
        double chi[6][24];
        for (int lc = 0; lc < 6; lc++) {
          double ua[24];
          affine_local(lc, ua);
          for (int v = 0; v < 8; v++) {
            const int n = (wrap(i + ((v >> 2) & 1), nx) * ny + wrap(j + ((v >> 1) & 1), ny)) * nz + wrap(k + (v & 1), nz);
            for (int c = 0; c < 3; c++) chi[lc][3 * v + c] = ua[3 * v + c] + u[((size_t)lc * 3 + c) * V + n];
          }
        }
        for (int q = 0; q < 8; q++) {
          const double xi[3] = {(q & 4) ? g : -g, (q & 2) ? g : -g, (q & 1) ? g : -g};
          double gN[8][3];
          for (int v = 0; v < 8; v++) {
            const int sgn[3] = {((v >> 2) & 1) ? 1 : -1, ((v >> 1) & 1) ? 1 : -1, (v & 1) ? 1 : -1};
            gN[v][0] = sgn[0] / 4.0 * (1 + sgn[1] * xi[1]) * (1 + sgn[2] * xi[2]);
            gN[v][1] = sgn[1] / 4.0 * (1 + sgn[0] * xi[0]) * (1 + sgn[2] * xi[2]);
            gN[v][2] = sgn[2] / 4.0 * (1 + sgn[0] * xi[0]) * (1 + sgn[1] * xi[1]);
          }
          double a[6][6];
          for (int lc = 0; lc < 6; lc++) {
            double gu[3][3] = {{0}};
            for (int v = 0; v < 8; v++)
              for (int c = 0; c < 3; c++)
                for (int d = 0; d < 3; d++) gu[c][d] += chi[lc][3 * v + c] * gN[v][d];
            a[0][lc] = gu[0][0];
            a[1][lc] = gu[1][1];
            a[2][lc] = gu[2][2];
            a[3][lc] = gu[1][2] + gu[2][1];
            a[4][lc] = gu[0][2] + gu[2][0];
            a[5][lc] = gu[0][1] + gu[1][0];
          }
          for (int p = 0; p < 6; p++)
            for (int s = 0; s < 6; s++) {
              double acc = 0;
              for (int r = 0; r < 6; r++)
                for (int t = 0; t < 6; t++) acc += a[r][p] * D[r * 6 + t] * a[t][s];
              M[p * 6 + s] += Ee * 0.125 * acc;
            }
        }
      }
  for (int p = 0; p < 6; p++)
    for (int s = 0; s < 6; s++) C[p * 6 + s] = M[p * 6 + s] / ((double)V * Tv[p] * Tv[s]);
}
