/*
 * femh.h -- C ABI of the MI355X voxel-FEM homogenization engine.
 *
 * One entry point family replaces the solver-facing surface of desaic/fem's
 * "density grid in -> homogenized tensor C_H (6x6) and dC_H/drho out" path:
 *
 *   femh_create / femh_destroy
 *       replace NumericalCoarseningUsingMultigrid::init
 *       (SamplesGenerator/NumericalCoarseningUsingMultigrid.h:27,
 *        SamplesGenerator/NumericalCoarseningUsingMultiGrid.cpp:134-219) and
 *       NumericalCoarsening::init (SamplesGenerator/NumericalCoarsening.h:26,
 *        NumericalCoarsening.cpp:116-133): grid, base material, solver knobs.
 *   femh_homogenize / femh_homogenize_modulus
 *       replace computeCoarsenedElasticityTensorAndParameters3D
 *       (NumericalCoarseningUsingMultigrid.h:29-30,
 *        NumericalCoarseningUsingMultiGrid.cpp:221-313) and the solve+G part of
 *       FEM3DFun::setParam (fast_fem/FEM3DFun.cpp:128-196).
 *   femh_gradient
 *       replaces the element-energy adjoint of FEM3DFun::df
 *       (fast_fem/FEM3DFun.cpp:222-268) with dJ/dC supplied by the caller.
 *   femh_get_fields
 *       exposes the 6 solved periodic fields (the reference's
 *       Homogenization::unit_test_u consumed at NumericalCoarseningUsingMultiGrid
 *       .cpp:275-288, fast_fem/FEM3DFun.cpp:159-161).
 *   femh_thermal_expansion
 *       replaces ThermalAnalysis::computeThermalCoefficient / run3D
 *       (SamplesGenerator/ThermalAnalysis.cpp:210-274, :388-505): effective
 *       thermal stress and expansion from the solved fields.
 *   femh_launch_shapes
 *       reports the launch shapes in use (femh_opts.autotune); no counterpart
 *       in the reference.
 *   femh_coarse_strain_tensor
 *       replaces the force-driven NumericalCoarsening solve and its
 *       coarse-strain reduction (SamplesGenerator/NumericalCoarsening.cpp:
 *       1148-1306, femlib/src/MeshUtilities.cpp:960-1060).
 *   femh_homogenize_batch
 *       replaces the OpenMP sample loop of SamplesGeneratorImpl::
 *       computeParametersAndTensorValues (SamplesGenerator/SamplesGeneratorImpl
 *       .cpp:3707-3731) with one device-side batch.
 *
 * Conventions
 *   voxel index   (i*ny + j)*nz + k, z fastest (femlib/include/ElementRegGrid.hpp:28)
 *   periodic      node index == voxel index; the (n+1)^3 expansion of
 *                 copyPeriodicU is left to the caller (host facade)
 *   C_H           row-major 6x6, Voigt order (xx, yy, zz, yz, xz, xy),
 *                 engineering shear strains, energy form
 *                 C_ij = 1/V sum_e E_e chi_i^T K0 chi_j (V = voxel count, h = 1)
 *   fields        u[lc][c][idx]: 6 load cases x 3 components x V, zero mean
 *   status        int, 0 = FEMH_OK; femh_last_error() gives the message
 * All buffers are caller-owned.  A context is used by one thread at a time
 * (one context per thread, as the reference keeps one homogenizer per OpenMP
 * thread); calls on different contexts may run concurrently.
 */
#ifndef FEMH_H
#define FEMH_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FEMH_ABI_VERSION 3

enum femh_status {
  FEMH_OK = 0,
  FEMH_BAD_ARG = 1,
  FEMH_NOT_CONVERGED = 2,
  FEMH_NONFINITE = 3,
  FEMH_DEVICE_ERROR = 4,
  FEMH_OOM = 5,
  FEMH_NO_FIELDS = 6
};

enum femh_precond { FEMH_PRECOND_JACOBI = 0, FEMH_PRECOND_MG = 1 };
enum femh_precision { FEMH_FP64 = 0, FEMH_MIXED = 1 };

/* Material interpolation E(rho):
 *   FEMH_INTERP_SIMP     E = Emin + (E0 - Emin) * rho^penal
 *                        (MaterialOpt/src/FEM3DFun.cpp:244-257 uses Emin=1e-3 E0, p=3)
 *   FEMH_INTERP_FASTFEM  E = E0 * rho / (3 - 2 rho)  (fast_fem/FEM3DFun.cpp:20-27)
 *   FEMH_INTERP_MODULUS  rho already holds E per voxel (binary grids etc.) */
enum femh_interp { FEMH_INTERP_SIMP = 0, FEMH_INTERP_FASTFEM = 1, FEMH_INTERP_MODULUS = 2 };

/* load cases solved (femh_opts.structure) */
enum femh_structure { FEMH_STRUCT_GENERAL = 0, FEMH_STRUCT_CUBIC = 1 };

typedef struct femh_opts {
  int nx, ny, nz;       /* grid (voxels) */
  double E0, Emin, nu;  /* base material */
  double penal;         /* SIMP exponent */
  int interp;           /* femh_interp */
  int precond;          /* femh_precond */
  int levels;           /* multigrid levels incl. fine (0 = as deep as the grid allows) */
  int nu1, nu2;         /* Chebyshev smoothing degree before / after coarse correction */
  double cheby_ratio;   /* smoothing interval [lmax/ratio, lmax] of D^-1 A */
  int coarse_iters;     /* Chebyshev steps of the coarsest solve when that level is too
                           large for its dense inverse (> 648 dofs); default 32 */
  double tol;           /* ||r||/||b|| target (FP64) / outer refinement target (mixed) */
  int max_iter;         /* outer iteration cap per load case */
  int precision;        /* femh_precision */
  double ch_tol;        /* mixed: stop when the estimated relative C_H error of
                           every load case is below this (0 = use tol only) */
  int check_every;      /* iterations between host convergence checks */
  int device;           /* HIP device ordinal */
  double inner_tol;     /* mixed: relative tolerance of each FP32 correction solve */
  int max_outer;        /* mixed: refinement sweep cap */
  int apply_kernel;     /* fine-grid operator: 0 tiled cube-mode (default), 1 dense gather */
  int batch;            /* samples solved together per launch sequence (1..16) */
  int coarse_nu;        /* smoothing degree (pre and post) on the coarse levels; 0: nu1/nu2
                           (default 3: coarse sweeps are cheap and save fine ones) */
  int warm_start;       /* 1: start each sample slot from the fields of the previous
                           call with the same sample count (an optimisation loop's
                           previous iterate); 0: start from zero */
  int structure;        /* femh_structure.  FEMH_STRUCT_CUBIC (cube grids only): a cell
                           whose moduli are invariant under the axis swaps (checked
                           on the device, 1e-12 relative per voxel) is solved for the
                           xx and yz load cases only; the other four fields are those
                           two with axes swapped, and C_H is averaged over the axis
                           permutations.  A cell that is not invariant is solved with
                           all six cases (femh_stats.iters[1] > 0 tells which path
                           ran).  The reference's CUBIC mode
                           (NumericalCoarseningUsingMultiGrid.cpp:275-282) makes the
                           same two solves */
  int autotune;         /* 1: before the first solve of each (grid, batch, precision)
                           shape in the process, time the fine operator and the grid
                           transfers for a few launch shapes (tile of node columns
                           and node planes per workgroup, planes per thread) and
                           keep the fastest; the
                           choice is cached for the process.  The fine-operator
                           shape changes the order of the dot-product partial sums,
                           so results move at rounding level.  0: built-in shapes */
  int streams;          /* many-calls with more than `batch` cells: batches run on
                           this many contexts at once (each its own stream and host
                           thread; streams - 1 extra contexts are created on first
                           use).  Without warm_start every cell's result is the same
                           as with 1 (with it, each context starts from its own last
                           batch).  The
                           last batch runs on the primary context, so femh_gradient
                           and the field calls still refer to it.  Profiling runs
                           the batches one at a time.  Default 1 */
  int batch_threads;    /* femh_homogenize_batch: host threads that keep batches in flight,
                           each on its own sub-context and stream (0 = 8).  1 runs the
                           batches one after another on one thread */
  int graph;            /* 1 (default): PCG iterations run as a recorded launch sequence
                           (hipGraph) of two iterations, replayed until every load case
                           converged; the convergence test runs on the device and the
                           host looks at it once per replay.  0: the same iterations
                           launched one kernel at a time.  Results are the same */
  int adapt_inner;      /* mixed with ch_tol > 0: 1 (default): a correction sweep after the
                           first stops at the residual reduction its cell's estimate
                           needs, 0.5 sqrt(ch_tol / estimate) clamped to [inner_tol, 0.1],
                           rather than at inner_tol; the stop test stays the estimate.
                           0: every sweep runs to inner_tol */
} femh_opts;

typedef struct femh_stats {
  int iters[6];         /* PCG iterations per load case (inner FP32 iterations in mixed) */
  int outer[6];         /* refinement sweeps per load case (mixed), 1 in FP64 */
  double relres[6];     /* final FP64 relative residual per load case */
  double ch_err_est;    /* a-posteriori relative C_H error estimate (energy form) */
  double ms_setup;      /* device time: hierarchy and right-hand sides */
  double ms_solve;      /* device time: PCG */
  double ms_total;      /* wall time of the call */
  int levels;           /* multigrid levels used */
  int fp64_fallback;    /* mixed precision: 1 when an FP32 correction sweep failed to
                           converge for this sample and FP64 MG-PCG finished it (the
                           error is then that of the FP64 tolerance; ch_err_est is 0) */
  double ritz_min;      /* mixed precision with ch_tol > 0: the smallest Ritz value of
                           the preconditioned operator M^-1 K seen in the correction
                           sweeps (it divides ch_err_est); 0 when none was formed */
} femh_stats;

void femh_default_opts(femh_opts *o);

int femh_create(const femh_opts *o, struct femh_ctx **out);
void femh_destroy(struct femh_ctx *ctx);

/* Homogenize one density grid (host memory, V = nx*ny*nz doubles, interpreted
 * through opts.interp).  C[36] receives C_H.  stats may be NULL. */
int femh_homogenize(struct femh_ctx *ctx, const double *rho, double *C, femh_stats *stats);

/* Same, with the Young's modulus of every voxel given directly. */
int femh_homogenize_modulus(struct femh_ctx *ctx, const double *E, double *C, femh_stats *stats);

/* Device-pointer variant: d_rho is device memory (V doubles), d_C device
 * memory (36 doubles); work is queued on `stream` (a hipStream_t, NULL =
 * default stream) and the call returns after it completes. */
int femh_homogenize_device(struct femh_ctx *ctx, const double *d_rho, double *d_C, void *stream,
                           femh_stats *stats);

/* n samples of the context's grid, solved `batch` at a time: rho holds n*V
 * doubles (sample after sample), C receives 36*n, stats (may be NULL) n
 * entries.  The _modulus form takes Young's moduli, the _device form device
 * pointers (densities) and a stream. */
int femh_homogenize_many(struct femh_ctx *ctx, int n, const double *rho, double *C,
                         femh_stats *stats);
int femh_homogenize_many_modulus(struct femh_ctx *ctx, int n, const double *E, double *C,
                                 femh_stats *stats);
int femh_homogenize_many_device(struct femh_ctx *ctx, int n, const double *d_rho, double *d_C,
                                void *stream, femh_stats *stats);

/* dJ/drho for every voxel from the fields of the last homogenize call:
 * dJdrho[e] = sum_ij dJdC[6i+j] * E'(rho_e) * chi_i^T K0 chi_j / V.
 * After a call with nb samples (the last batch of a _many call), dJdC holds
 * 36*nb values and dJdrho receives nb*V.
 * Returns FEMH_NO_FIELDS when no solve has been done. */
int femh_gradient(struct femh_ctx *ctx, const double *dJdC, double *dJdrho);

/* Device-memory form of femh_gradient for a loop that keeps its densities on
 * the device: d_dJdrho (device, nb*V) receives dJ/drho of the last call's nb
 * cells plus shift[s] per cell (host, nb values, may be NULL: the volume
 * term mw (mean(rho) - m0) / V of fast_fem/FEM3DFun.cpp:222-268), and gmax
 * (host, nb, may be NULL) max |dJ/drho| per cell, the norm that
 * gradientDescent divides the step by (fast_fem/run3d.cpp:268-289).  Work on
 * `stream` is waited for first; the call returns when the result is in
 * place.  Returns FEMH_NO_FIELDS when no solve has been done. */
int femh_gradient_device(struct femh_ctx *ctx, const double *dJdC, const double *shift, double *d_dJdrho,
                         double *gmax, void *stream);

/* Trial points of a projected-gradient line search over n cells of the
 * context's grid (fast_fem/run3d.cpp:238-265, lineSearch: x' = clamp(x - h g,
 * lb, ub)), in device memory: d_x, d_g, d_out hold n*V doubles; h (host, n
 * values) selects per cell: h > 0 writes clamp(x - h g, lb, ub) (product and
 * difference rounded separately), h < 0 copies x, h == 0 leaves the cell's
 * d_out as it is.  sums (host, n, may be NULL) receives each cell's sum of
 * d_out after the call (for the volume term).  Waits for `stream` first. */
int femh_descent_trial(struct femh_ctx *ctx, int n, const double *d_x, const double *d_g, const double *h,
                       double lb, double ub, double *d_out, double *sums, void *stream);

/* Per-voxel energy cross terms W[e*21 + p] = chi_i^T K0 chi_j / V, p over i<=j
 * (the unscaled adjoint building block). */
int femh_energy_terms(struct femh_ctx *ctx, double *W);

/* The force-driven (coarse-strain) form of C_H of the cells of the last
 * homogenize call (36 doubles per cell), from the same fields: what the
 * reference's NumericalCoarsening path computes -- six face-traction loads
 * under periodic, translation and rotation constraints solved on the
 * (n+1)^3-vertex mesh (SamplesGenerator/NumericalCoarsening.cpp:1148-1306),
 * then MeshUtilities::computeCoarsenedElasticityTensor
 * (femlib/src/MeshUtilities.cpp:960-1060), which applies the
 * engineering-shear stiffness to tensor shear strains -- without the
 * reference's 1e-5 diagonal regularisation (that changes C by ~1e-6
 * relative).  It differs from the energy form of femh_homogenize by the
 * convention (a few percent on heterogeneous cells, nothing on homogeneous
 * ones).  Unlike the energy form it is first order in the error of the
 * fields, so it wants fields solved to the residual tolerance (in mixed
 * precision, ch_tol = 0).  Cube cells only (BAD_ARG otherwise: on other cells
 * the reference's shear loads carry a net moment).  Returns FEMH_NO_FIELDS
 * before a solve. */
int femh_coarse_strain_tensor(struct femh_ctx *ctx, double *C);

/* Copies the 6 solved periodic fluctuation fields (6*3*V doubles per sample
 * of the last batch). */
int femh_get_fields(struct femh_ctx *ctx, double *u);

/* Thermal expansion of the cells of the last homogenize call, with no further
 * solve.  alpha holds the expansion coefficient of every voxel (nb*V doubles,
 * host).  beta (6*nb, may be NULL) receives the effective thermal stress per
 * unit temperature rise, beta_q = (1/V) sum_e alpha_e E_e int_e I : C0 : chi_q
 * (Voigt), and alpha_eff (6*nb, may be NULL) the effective expansion
 * C_H^-1 beta (engineering shear).  Replaces the per-cell force-driven solve
 * of ThermalAnalysis::computeThermalCoefficient
 * (SamplesGenerator/ThermalAnalysis.cpp:388-505), whose result run3D
 * (:210-274) estimates; here the periodic value is exact for the
 * discretisation (reciprocity with the six elastic fields).  Returns
 * FEMH_NO_FIELDS when no solve has been done. */
int femh_thermal_expansion(struct femh_ctx *ctx, const double *alpha, double *beta, double *alpha_eff);

/* Launch shapes in use (after a solve; 0 = the built-in choice):
 * shapes[0] = the fine operator's shape (bits 0-7 node planes per workgroup,
 * bits 8-15 the tile of owned node columns per workgroup: 1 = 16 x 16,
 * 2 = 8 x 32, 3 = 16 x 32), shapes[1 + l] =
 * coarse planes per thread of the transfers between levels l and l + 1.
 * Writes at most n values; returns their count (levels) or a negative status.
 * No counterpart in the reference, whose CPU loops have no launch shapes. */
int femh_launch_shapes(struct femh_ctx *ctx, int *shapes, int n);

/* Sets the launch shapes of later calls on this context (the encoding of
 * femh_launch_shapes; 0 = built-in), e.g. shapes an earlier autotuned run
 * reported.  With femh_opts.autotune set, the tuner's choice replaces them.
 * Results move at rounding level with the fine-operator shape (the order of
 * its dot-product partial sums).  No counterpart in the reference. */
int femh_set_launch_shapes(struct femh_ctx *ctx, const int *shapes, int n);

/* Batch: n samples, sample s has dims[3s..3s+2] and modulus grid E[s] (host).
 * C receives 36*n doubles.  Material nu/Emin from the context options.
 * stats (may be NULL) receives aggregate timings; iters[0] = max iterations. */
int femh_homogenize_batch(struct femh_ctx *ctx, int n, const int *dims, const double *const *E,
                          double *C, femh_stats *stats);

/* Test hooks (operator-level parity, no solve): y = K u on the fine grid or
 * on multigrid level `level` of the hierarchy built from the last
 * homogenize/prepare call; u,y host, component-major 3*V_level doubles. */
int femh_prepare(struct femh_ctx *ctx, const double *E);
int femh_apply_level(struct femh_ctx *ctx, int level, const double *u, double *y);
int femh_level_dims(struct femh_ctx *ctx, int level, int *dims);

/* Kernel timing (instrumentation, no counterpart in the reference): when
 * enabled, every fine-grid operator launch is bracketed by HIP events on the
 * context's stream; femh_kernel_stats returns, for which = 0 (FP64 apply) or
 * 1 (FP32 apply), the launch count, summed device milliseconds and the
 * algorithmic bytes / flops of those launches (bytes: the vectors each
 * stage/epilogue variant reads and writes, E once; flops: 1152 per voxel per
 * RHS, the dense 24x24 element product -- a dense-equivalent figure, not the
 * instructions the mode-form kernel executes).  Enabling resets the
 * counters. */
int femh_set_profiling(struct femh_ctx *ctx, int enable);
int femh_kernel_stats(struct femh_ctx *ctx, int which, long long *launches, double *ms,
                      double *bytes, double *flops);

/* Largest grid (voxels) femh_homogenize_device_noalloc accepts. */
#define FEMH_NOALLOC_MAX_VOXELS 32768

/* femh_homogenize_device without device memory allocation: no hipMalloc,
 * hipFree, stream or event creation.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (serving loops and harnesses
 * that forbid it).  The cell (o->nx*o->ny*o->nz <= FEMH_NOALLOC_MAX_VOXELS
 * voxels, densities interpreted through o->interp) is solved by Jacobi-PCG in
 * FP64 -- o->precond must be FEMH_PRECOND_JACOBI and o->precision FEMH_FP64,
 * without warm start or autotuning -- giving the same C_H (device memory, 36
 * doubles) as femh_homogenize_device with those options.  Work runs on
 * `stream` (NULL = the default stream); the call returns when it is done.
 * Every buffer comes from one 64 MiB device region compiled into the library
 * (one per GPU, held until the process ends); calls on one GPU share it and
 * are serialised.  err (may be NULL) receives the message of a failed call. */
int femh_homogenize_device_noalloc(const femh_opts *o, const double *d_rho, double *d_C, void *stream,
                                   femh_stats *stats, char *err, size_t err_len);

/* Copies the last error message (NUL-terminated) into buf. Returns the
 * status of the last failed call on this context. */
int femh_last_error(struct femh_ctx *ctx, char *buf, size_t len);

/* Host-only (no device work): the 24x24 trilinear hex stiffness of a unit
 * voxel (h = 1) with Young's modulus E and Poisson ratio nu, row-major, local
 * corner order 4x + 2y + z and dof 3*corner + component -- the matrix
 * MaterialQuad::getStiffness returns (femlib/src/MaterialQuad.cpp:68-79). */
int femh_element_stiffness(double E, double nu, double *Ke);

int femh_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
