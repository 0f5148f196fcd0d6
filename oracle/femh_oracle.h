/*
 * femh_oracle.h -- CPU FP64 restatement of desaic/fem's periodic voxel
 * homogenization path.  TEST INFRASTRUCTURE ONLY: used by tests/, by the
 * smoke() check and by bench.py's cpu_baseline leg as the checker; never by
 * the product library.
 *
 * Conventions (shared with the product, stated once here):
 *   grid       nx*ny*nz voxels = elements = periodic nodes, unit voxel h=1
 *   index      idx(i,j,k) = (i*ny + j)*nz + k          (z fastest)
 *              reference: femlib/include/ElementRegGrid.hpp:28 (elements),
 *              SamplesGenerator/NumericalCoarseningUsingMultiGrid.cpp:16-41
 *              (periodic nodes)
 *   corners    local v = 4*bx + 2*by + bz              (z fastest)
 *              reference: femlib/src/ElementHex.cpp:24-33 (sw table),
 *              NumericalCoarseningUsingMultiGrid.cpp:72-74 (oneV)
 *   dofs       element-local 3*v + c; global fields are component-major
 *              u[c*V + idx]
 *   load cases Voigt order (xx, yy, zz, yz, xz, xy), engineering shear
 *              (unit case 3 means eps_yz = eps_zy = 1/2)
 */
#ifndef FEMH_ORACLE_H
#define FEMH_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* 24x24 trilinear hex stiffness, row-major, for Young's modulus E, Poisson
 * nu, cube edge h.  float_gauss!=0 rounds the Gauss abscissa 1/sqrt(3) to
 * float the way femlib/src/Quadrature.cpp:3-7 does. */
void or_ke_hex(double E, double nu, double h, int float_gauss, double *Ke);

/* Lame parameters, Utilities/src/cfgMaterialUtilities.cpp:2945-2955. */
void or_lame(double E, double nu, double *lambda, double *mu);

/* Unit macro strain of load case lc as a symmetric 3x3 tensor. */
void or_macro_strain(int lc, double eps[9]);

/* y = K u on the periodic grid (one right-hand side, component-major). */
void or_apply(int nx, int ny, int nz, const double *E, const double *Ke,
              const double *u, double *y);
/* point diagonal of K, component-major [3*V] */
void or_diag(int nx, int ny, int nz, const double *E, const double *Ke,
             double *d);
/* b = -K u0 for the affine field of load case lc */
void or_rhs(int nx, int ny, int nz, const double *E, const double *Ke,
            int lc, double *b);

/* Preconditioned CG on the periodic system for one right-hand side.
 * precond: 0 = Jacobi, 1 = geometric multigrid (Galerkin, V(nu1,nu2)).
 * Stops when ||r||_2 <= tol*||b||_2.  Returns 0 converged, 1 not converged.
 * x is returned with zero mean per component. */
typedef struct {
  int precond;      /* 0 Jacobi, 1 multigrid */
  int levels;       /* multigrid levels (including fine) */
  int nu1, nu2;     /* pre/post Jacobi sweeps */
  double omega;     /* Jacobi damping (times 1/lambda_max) for smoother 0 */
  int smoother;     /* 0 damped Jacobi, 1 Chebyshev */
  double cheby_ratio; /* Chebyshev interval [lmax/ratio, lmax] */
  int coarse_iters; /* inner Jacobi-PCG iterations when coarsest is large */
  double tol;
  int max_iter;
} or_solver_opts;

void or_default_opts(or_solver_opts *o);

/* Full homogenization: solves the 6 load cases, returns C[36] (row-major,
 * Voigt, engineering shear), the fluctuation fields u[6][3*V] (zero mean),
 * per-case iteration counts and final relative residuals. */
int or_homogenize(int nx, int ny, int nz, const double *E, double nu,
                  const or_solver_opts *o, double *C, double *u,
                  int *iters, double *relres);

/* Same, for the load cases set in case_mask (bit lc; the others return zero
 * fields; C is written only when all six are set), with wall seconds of the
 * set-up, of the solves and of C_H in seconds[3] (may be NULL) -- the CPU
 * baseline times a bounded sample with it. */
int or_homogenize_cases(int nx, int ny, int nz, const double *E, double nu,
                        const or_solver_opts *o, int case_mask, double *C, double *u,
                        int *iters, double *relres, double *seconds);

/* Energy-form C_H from fluctuation fields (NumericalCoarseningUsingMultiGrid
 * .cpp:290-313, with the affine part re-added element-locally). */
void or_ch_energy(int nx, int ny, int nz, const double *E, double nu,
                  const double *u, double *C);

/* Per-voxel element energy cross terms W_e[ij] = chi_i^T Ke(E=1) chi_j / V
 * (i<=j packed 21 per voxel: W[21*e + p]) -- the adjoint building block
 * (fast_fem/FEM3DFun.cpp:222-268). */
void or_energy_terms(int nx, int ny, int nz, double nu, const double *u,
                     double *W);

/* dJ/dE_e = sum_ij G[6i+j] * chi_i^T Ke chi_j / V  (G = dJ/dC full 6x6). */
/* beta[6] = (1/V) sum_e alpha_e E_e int_e I : C0 : eps(affine_q + u_q) (the
 * effective thermal stress per unit temperature rise; u as or_homogenize) */
void or_thermal_stress(int nx, int ny, int nz, const double *E, double nu, const double *alpha,
                       const double *u, double *beta);
void or_grad_E(int nx, int ny, int nz, double nu, const double *u,
               const double *G, double *dJdE);

/* Multigrid pieces exposed for tests: coarse operator apply at level l of
 * the Galerkin hierarchy built from E (element-matrix form). */
void *or_mg_create(int nx, int ny, int nz, const double *E, double nu,
                   int levels);
void or_mg_destroy(void *mg);
int or_mg_level_dims(void *mg, int l, int *dims);
void or_mg_apply(void *mg, int l, const double *u, double *y);
void or_mg_restrict(void *mg, int l, const double *rf, double *rc);
void or_mg_prolong_add(void *mg, int l, const double *ec, double *ef);

/* Force-driven periodic homogenization of the reference's NumericalCoarsening
 * path on the (n+1)^3-vertex mesh (small grids: dense KKT solve).  C[36]
 * receives the coarse-strain tensor (Voigt, engineering shear), u[6][3 NV]
 * the vertex displacements of the six face-traction load cases (reference
 * order xx, yy, zz, xy, yz, xz; vertex (i*(ny+1)+j)*(nz+1)+k, dof 3 v + c).
 * regularize: diagonal x (1 + 1e-5) as the reference assembles it;
 * faithful: the reference's float roundings (Gauss abscissa, vertex spacing
 * 1/max n, the 1/nelem factor).  Returns 0 on success. */
int or_force_driven(int nx, int ny, int nz, const double *E, double nu, int regularize,
                    int faithful, double *C, double *u);
/* the constraint rows of that system, dense [nrows][3 NV]; with rows == NULL
 * returns the row count */
int or_kkt_constraints(int nx, int ny, int nz, int faithful, double *rows, int max_rows);
/* the coarse-strain tensor of given vertex displacement fields */
void or_coarse_strain_tensor(int nx, int ny, int nz, const double *E, double nu, int faithful,
                             const double *u, double *C);
/* the same tensor from strain-driven periodic fields u[6][3][V] */
void or_coarse_strain_from_fields(int nx, int ny, int nz, const double *E, double nu, int float_gauss,
                                  const double *u, double *C);

double or_wall(void);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
