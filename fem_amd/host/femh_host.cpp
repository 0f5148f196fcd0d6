// femh_host.cpp -- C++ host facade over the femh C ABI (see femh_host.hpp).
#include "femh_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>

namespace femh {

namespace {

[[noreturn]] void fail(femh_ctx *ctx, int st, const char *what) {
  char buf[512] = {0};
  if (ctx) femh_last_error(ctx, buf, sizeof(buf));
  throw HostError(st, std::string(what) + ": " + buf);
}

void check(femh_ctx *ctx, int st, const char *what) {
  if (st != FEMH_OK) fail(ctx, st, what);
}

// 6x6 inverse by Gauss-Jordan with partial pivoting
Mat6 inverse6(const Mat6 &A) {
  double M[6][12];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 12; j++) M[i][j] = j < 6 ? A(i, j) : (j - 6 == i ? 1.0 : 0.0);
  for (int c = 0; c < 6; c++) {
    int p = c;
    for (int r = c + 1; r < 6; r++)
      if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
    if (M[p][c] == 0.0) throw HostError(FEMH_BAD_ARG, "singular 6x6 matrix");
    if (p != c)
      for (int j = 0; j < 12; j++) std::swap(M[p][j], M[c][j]);
    const double inv = 1.0 / M[c][c];
    for (int j = 0; j < 12; j++) M[c][j] *= inv;
    for (int r = 0; r < 6; r++) {
      if (r == c) continue;
      const double f = M[r][c];
      for (int j = 0; j < 12; j++) M[r][j] -= f * M[c][j];
    }
  }
  Mat6 out;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) out(i, j) = M[i][j + 6];
  return out;
}

}  // namespace

void BaseMaterial::youngPoisson(double &E, double &nu) const {
  E = mu * (3 * lambda + 2 * mu) / (lambda + mu);
  nu = lambda / (2 * (lambda + mu));
}

VectorXd RealFun::df() { throw HostError(FEMH_BAD_ARG, "df() is not implemented by this function"); }

// ---- fields ----------------------------------------------------------------
void PiecewiseConstant3D::allocate(int nx, int ny, int nz) {
  gridSize[0] = nx;
  gridSize[1] = ny;
  gridSize[2] = nz;
  param.assign((size_t)nx * ny * nz, 0.0);
}

void PiecewiseConstant3D::gridIdx(const double x[3], int idx[3]) const {
  for (int d = 0; d < 3; d++) {
    int i = (int)(x[d] * gridSize[d]);
    idx[d] = std::min(std::max(i, 0), gridSize[d] - 1);
  }
}

void PiecewiseConstantSym3D::gridIdx(const double x[3], int idx[3]) const {
  for (int d = 0; d < 3; d++) {
    double c = 2 * x[d];
    if (c > 1) c = 2 - c;
    int i = (int)(c * gridSize[d]);
    idx[d] = std::min(std::max(i, 0), gridSize[d] - 1);
  }
}

void PiecewiseConstantCubic3D::gridIdx(const double x[3], int idx[3]) const {
  double c[3];
  for (int d = 0; d < 3; d++) {
    c[d] = 2 * x[d];
    if (c[d] > 1) c[d] = 2 - c[d];
  }
  // sort descending: the parameter lives on one of the 48 symmetric copies
  if (c[0] < c[1]) std::swap(c[0], c[1]);
  if (c[0] < c[2]) std::swap(c[0], c[2]);
  if (c[1] < c[2]) std::swap(c[1], c[2]);
  for (int d = 0; d < 3; d++) {
    int i = (int)(c[d] * gridSize[d]);
    idx[d] = std::min(std::max(i, 0), gridSize[d] - 1);
  }
}

double PiecewiseConstant3D::f(const double x[3]) {
  int idx[3];
  gridIdx(x, idx);
  return param[((size_t)idx[0] * gridSize[1] + idx[1]) * gridSize[2] + idx[2]];
}

std::vector<std::pair<int, double>> PiecewiseConstant3D::df(const double x[3]) {
  int idx[3];
  gridIdx(x, idx);
  return {{(int)(((size_t)idx[0] * gridSize[1] + idx[1]) * gridSize[2] + idx[2]), 1.0}};
}

// ---- FEM3DFun --------------------------------------------------------------
FEM3DFun::FEM3DFun() : gridSize(3, 0) { femh_default_opts(&solverOpts); }

FEM3DFun::~FEM3DFun() {
  if (ctx_) femh_destroy(ctx_);
}

void FEM3DFun::initArrays() {
  G = Mat6{};
  G0 = Mat6{};
  wG = Mat6{};
}

void FEM3DFun::ensureContext() {
  if (ctx_) return;
  if (gridSize.size() != 3 || gridSize[0] < 2 || gridSize[1] < 2 || gridSize[2] < 2)
    throw HostError(FEMH_BAD_ARG, "FEM3DFun: gridSize must hold three sizes >= 2");
  femh_opts o = solverOpts;
  o.nx = gridSize[0];
  o.ny = gridSize[1];
  o.nz = gridSize[2];
  o.E0 = 1.0;
  o.nu = nu;
  o.interp = FEMH_INTERP_FASTFEM;
  int st = femh_create(&o, &ctx_);
  if (st != FEMH_OK) {
    femh_ctx *c = ctx_;
    ctx_ = nullptr;
    char buf[512] = {0};
    if (c) {
      femh_last_error(c, buf, sizeof(buf));
      femh_destroy(c);
    }
    throw HostError(st, std::string("femh_create: ") + buf);
  }
}

void FEM3DFun::init(const VectorXd &x0) {
  param = x0;
  ensureContext();
}

void FEM3DFun::setParam(const VectorXd &x0) {
  if (!field) throw HostError(FEMH_BAD_ARG, "FEM3DFun: no field");
  ensureContext();
  param = x0;
  for (size_t i = 0; i < field->param.size() && i < x0.size(); i++) field->setParam((int)i, x0[i]);
  const int nx = gridSize[0], ny = gridSize[1], nz = gridSize[2];
  distribution.assign((size_t)nx * ny * nz, 0.0);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const double x[3] = {(i + 0.5) / nx, (j + 0.5) / ny, (k + 0.5) / nz};
        const double v = field->f(x);
        distribution[((size_t)i * ny + j) * nz + k] = std::min(std::max(v, 1e-3), 1.0);
      }
  Mat6 C;
  check(ctx_, femh_homogenize(ctx_, distribution.data(), C.a.data(), &stats_), "femh_homogenize");
  // the reference accumulates G(i, j) for j >= i only (fast_fem/FEM3DFun.cpp:165-175)
  G = Mat6{};
  for (int a = 0; a < 6; a++)
    for (int b = a; b < 6; b++) G(a, b) = C(a, b);
  density = 0;
  for (double d : distribution) density += d;
  density /= (double)distribution.size();
  const double dx = 1.0 / (double)std::max(nx, std::max(ny, nz));
  Ke0.assign(1, VectorXd(576));
  check(nullptr, femh_element_stiffness(1.0, nu, Ke0[0].data()), "femh_element_stiffness");
  for (double &v : Ke0[0]) v *= dx;  // 3D element stiffness scales with the cell size
  if (!keepDisplacements) return;
  // copyPeriodicU: the fluctuation repeats on the wrapped nodes and the
  // affine part continues across them
  const size_t V = (size_t)nx * ny * nz;
  VectorXd chi(18 * V);
  check(ctx_, femh_get_fields(ctx_, chi.data()), "femh_get_fields");
  static const double eps[6][3][3] = {
      {{1, 0, 0}, {0, 0, 0}, {0, 0, 0}},     {{0, 0, 0}, {0, 1, 0}, {0, 0, 0}},
      {{0, 0, 0}, {0, 0, 0}, {0, 0, 1}},     {{0, 0, 0}, {0, 0, .5}, {0, .5, 0}},
      {{0, 0, .5}, {0, 0, 0}, {.5, 0, 0}},   {{0, .5, 0}, {.5, 0, 0}, {0, 0, 0}}};
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  u.assign(6, VectorXd((size_t)3 * Nx * Ny * Nz));
  for (int lc = 0; lc < 6; lc++)
    for (int i = 0; i < Nx; i++)
      for (int j = 0; j < Ny; j++)
        for (int k = 0; k < Nz; k++) {
          const size_t m = ((size_t)(i % nx) * ny + (j % ny)) * nz + (k % nz);
          const double X[3] = {(double)i, (double)j, (double)k};
          for (int d = 0; d < 3; d++) {
            const double aff = eps[lc][d][0] * X[0] + eps[lc][d][1] * X[1] + eps[lc][d][2] * X[2];
            u[lc][3 * (((size_t)i * Ny + j) * Nz + k) + d] = dx * (aff + chi[(size_t)(3 * lc + d) * V + m]);
          }
        }
}

double FEM3DFun::f() {
  double sum = 0;
  for (int q = 0; q < 36; q++) {
    const double d = G.a[q] - G0.a[q];
    sum += wG.a[q] * d * d;
  }
  return 0.5 * sum + 0.5 * mw * (density - m0) * (density - m0);
}

VectorXd FEM3DFun::df() {
  ensureContext();
  const size_t nE = distribution.size();
  // dJ/dG(i, j) for every (i, j), as fast_fem/FEM3DFun.cpp:236-249 sums them
  double dJdC[36];
  for (int q = 0; q < 36; q++) dJdC[q] = wG.a[q] * (G.a[q] - G0.a[q]);
  VectorXd dfddist(nE);
  check(ctx_, femh_gradient(ctx_, dJdC, dfddist.data()), "femh_gradient");
  const double dm = (mw / (double)nE) * (density - m0);
  for (double &v : dfddist) v += dm;
  VectorXd grad(param.size(), 0.0);
  const int nx = gridSize[0], ny = gridSize[1], nz = gridSize[2];
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const double x[3] = {(i + 0.5) / nx, (j + 0.5) / ny, (k + 0.5) / nz};
        const size_t e = ((size_t)i * ny + j) * nz + k;
        for (const auto &iv : field->df(x)) grad[iv.first] += dfddist[e] * iv.second;
      }
  return grad;
}

VectorXd FEM3DFun::fields() {
  ensureContext();
  VectorXd u((size_t)18 * gridSize[0] * gridSize[1] * gridSize[2]);
  check(ctx_, femh_get_fields(ctx_, u.data()), "femh_get_fields");
  return u;
}

void FEM3DFun::log(std::ostream &out) {
  for (int i = 0; i < 6; i++) {
    for (int j = 0; j < 6; j++) out << G(i, j) << (j < 5 ? " " : "\n");
  }
}

// ---- optimisation helpers (fast_fem/run3d.cpp:238-300) ---------------------
double infNorm(const VectorXd &a) {
  double m = 0;
  for (double v : a) m = std::max(m, std::fabs(v));
  return m;
}

void clampVector(VectorXd &x, const VectorXd &lb, const VectorXd &ub) {
  for (size_t i = 0; i < x.size(); i++) {
    if (i < lb.size()) x[i] = std::max(x[i], lb[i]);
    if (i < ub.size()) x[i] = std::min(x[i], ub[i]);
  }
}

int lineSearch(RealFun *fun, VectorXd &x0, const VectorXd &dir, double &h) {
  const double f0 = fun->f();
  for (int step = 0; step < 10; step++) {
    VectorXd x(x0.size());
    for (size_t i = 0; i < x.size(); i++) x[i] = x0[i] - h * dir[i];
    clampVector(x, fun->lowerBounds, fun->upperBounds);
    fun->setParam(x);
    if (fun->f() < f0) {
      x0 = x;
      return 0;
    }
    h /= 2;
  }
  return -1;
}

void gradientDescent(RealFun *fun, VectorXd &x0, int nSteps, double maxStep, int logInterval,
                     std::ostream *logfile) {
  VectorXd x = x0;
  fun->setParam(x);
  for (int ii = 0; ii < nSteps; ii++) {
    const VectorXd grad = fun->df();
    const double norm = infNorm(grad);
    if (!(norm > 0)) break;
    double h = maxStep / norm;
    if (lineSearch(fun, x, grad, h) < 0) break;
    x0 = x;
    if (logfile && logInterval > 0 && ii % logInterval == 0) fun->log(*logfile);
  }
  if (logfile) fun->log(*logfile);
}

double check_df(RealFun *fun, const VectorXd &x0, double h) {
  fun->setParam(x0);
  const VectorXd ana = fun->df();
  double max_diff = 0, max_val = 0;
  for (size_t i = 0; i < x0.size(); i++) {
    VectorXd x = x0;
    x[i] = x0[i] + h;
    fun->setParam(x);
    const double fp = fun->f();
    x[i] = x0[i] - h;
    fun->setParam(x);
    const double fm = fun->f();
    const double num = (fp - fm) / (2 * h);
    max_diff = std::max(max_diff, std::fabs(num - ana[i]));
    max_val = std::max(max_val, std::fabs(ana[i]));
  }
  fun->setParam(x0);
  return max_val > 0 ? max_diff / max_val : max_diff;
}

// ---- coarsening --------------------------------------------------------------
void repMaterialAssignment(int nx, int ny, int nz, const std::vector<int> &in, int repX, int repY, int repZ,
                           int sub, std::vector<int> &out) {
  const int NX = repX * nx * sub, NY = repY * ny * sub, NZ = repZ * nz * sub;
  out.assign((size_t)NX * NY * NZ, -1);
  for (int X = 0; X < NX; X++)
    for (int Y = 0; Y < NY; Y++)
      for (int Z = 0; Z < NZ; Z++) {
        const int m = (X / sub) % nx, n = (Y / sub) % ny, p = (Z / sub) % nz;
        out[((size_t)X * NY + Y) * NZ + Z] = in[((size_t)m * ny + n) * nz + p];
      }
}

double computeMaterialDensity(const std::vector<int> &a, const std::vector<double> &dens) {
  double r = 0;
  for (int m : a) r += dens.at(m);
  return a.empty() ? 0.0 : r / (double)a.size();
}

void cubicParameters(const Mat6 &C, double &E, double &nu, double &G) {
  G = C(3, 3);
  const double a = C(0, 1) / C(0, 0);
  nu = a / (1 + a);
  E = C(0, 1) * (1 + nu) * (1 - 2 * nu) / nu;
}

NumericalCoarseningUsingMultigrid::NumericalCoarseningUsingMultigrid() { femh_default_opts(&m_opts); }

NumericalCoarseningUsingMultigrid::~NumericalCoarseningUsingMultigrid() {
  if (m_ctx) femh_destroy(m_ctx);
  if (m_batchCtx) femh_destroy(m_batchCtx);
}

void NumericalCoarseningUsingMultigrid::init(int N[3], int rep, int sub, const std::vector<BaseMaterial> &mats,
                                             StructureType iType) {
  if (mats.size() < 2) throw HostError(FEMH_BAD_ARG, "need two base materials (void, solid)");
  m_nbBlockRep = rep;
  m_nbSubdiv = sub;
  m_baseMaterials = mats;
  for (int d = 0; d < 3; d++) m_n[d] = N[d] * rep * sub;
  // (NumericalCoarseningUsingMultiGrid.cpp:211: the structure type of init
  // picks the solve)
  const bool cube = m_n[0] == m_n[1] && m_n[1] == m_n[2];
  m_structure = iType == Cubic && cube ? FEMH_STRUCT_CUBIC : m_opts.structure;
  if (m_ctx) femh_destroy(m_ctx);
  m_ctx = nullptr;
  if (m_batchCtx) femh_destroy(m_batchCtx);
  m_batchCtx = nullptr;
  m_batch = 0;
  // the solid (material 1) is the base material; voxels of material 0 get
  // the coefficient m_densityMin (NumericalCoarseningUsingMultiGrid.cpp:186-193, :239)
  double E, nu;
  mats[1].youngPoisson(E, nu);
  femh_opts o = m_opts;
  o.nx = m_n[0];
  o.ny = m_n[1];
  o.nz = m_n[2];
  o.E0 = E;
  o.nu = nu;
  o.interp = FEMH_INTERP_MODULUS;
  o.structure = m_structure;
  const int st = femh_create(&o, &m_ctx);
  if (st != FEMH_OK) {
    char buf[512] = {0};
    if (m_ctx) {
      femh_last_error(m_ctx, buf, sizeof(buf));
      femh_destroy(m_ctx);
    }
    m_ctx = nullptr;
    throw HostError(st, std::string("femh_create: ") + buf);
  }
}

// moduli of one assignment on the engine grid (material 1 solid, others at
// m_densityMin, or void without variable density)
void NumericalCoarseningUsingMultigrid::moduli(const std::vector<int> &iMaterials, int N[3], double *oE) const {
  std::vector<int> cells;
  repMaterialAssignment(N[0], N[1], N[2], iMaterials, m_nbBlockRep, m_nbBlockRep, m_nbBlockRep, m_nbSubdiv,
                        cells);
  if ((int)cells.size() != m_n[0] * m_n[1] * m_n[2])
    throw HostError(FEMH_BAD_ARG, "material grid does not match init()");
  double E0, nu;
  m_baseMaterials[1].youngPoisson(E0, nu);
  for (size_t e = 0; e < cells.size(); e++)
    oE[e] = E0 * (m_useVariableDensity ? (cells[e] == 1 ? 1.0 : m_densityMin) : (cells[e] == 1 ? 1.0 : 0.0));
}

void NumericalCoarseningUsingMultigrid::computeCoarsenedElasticityTensorAndParameters3D(
    const std::vector<int> &iMaterials, int N[3], const std::vector<double> &dens, StructureType iType,
    std::vector<double> &ioTensorValues, std::vector<double> &ioParameters) {
  if (!m_ctx) throw HostError(FEMH_BAD_ARG, "init() was not called");
  std::vector<double> Emod((size_t)m_n[0] * m_n[1] * m_n[2]);
  moduli(iMaterials, N, Emod.data());
  femh_stats st;
  check(m_ctx, femh_homogenize_modulus(m_ctx, Emod.data(), m_C.a.data(), &st), "femh_homogenize_modulus");
  for (int i = 0; i < 6; i++)
    for (int j = 0; j <= i; j++) ioTensorValues.push_back(m_referenceTriangle && j < i ? 0.0 : m_C(i, j));
  if (iType == Cubic) {
    ioParameters.push_back(computeMaterialDensity(iMaterials, dens));
    double E, nuc, G;
    cubicParameters(m_C, E, nuc, G);
    ioParameters.push_back(E);
    ioParameters.push_back(nuc);
    ioParameters.push_back(G);
  }
}

void NumericalCoarseningUsingMultigrid::computeCoarsenedElasticityTensorAndParameters3D(
    const std::vector<std::vector<int>> &assign, int N[3], const std::vector<double> &dens, StructureType iType,
    std::vector<float> &oTensorValues, std::vector<float> &oParameters, int batch, int streams) {
  if (!m_ctx) throw HostError(FEMH_BAD_ARG, "init() was not called");
  if (batch < 1 || batch > 16) throw HostError(FEMH_BAD_ARG, "batch must be 1..16");
  if (streams < 1) throw HostError(FEMH_BAD_ARG, "streams must be at least 1");
  if (!m_batchCtx || m_batch != batch || m_streams != streams) {
    if (m_batchCtx) femh_destroy(m_batchCtx);
    m_batchCtx = nullptr;
    double E, nu;
    m_baseMaterials[1].youngPoisson(E, nu);
    femh_opts o = m_opts;
    o.nx = m_n[0];
    o.ny = m_n[1];
    o.nz = m_n[2];
    o.E0 = E;
    o.nu = nu;
    o.interp = FEMH_INTERP_MODULUS;
    o.structure = m_structure;
    o.batch = batch;
    o.streams = streams;
    const int st = femh_create(&o, &m_batchCtx);
    if (st != FEMH_OK) {
      char buf[512] = {0};
      if (m_batchCtx) {
        femh_last_error(m_batchCtx, buf, sizeof(buf));
        femh_destroy(m_batchCtx);
      }
      m_batchCtx = nullptr;
      throw HostError(st, std::string("femh_create: ") + buf);
    }
    m_batch = batch;
    m_streams = streams;
  }
  const size_t V = (size_t)m_n[0] * m_n[1] * m_n[2];
  // one engine call per `streams` batches, so that they are in flight together
  const size_t chunk = (size_t)batch * streams;
  std::vector<double> Emod(V * chunk), C(36 * chunk);
  std::vector<femh_stats> st(chunk);
  for (size_t c0 = 0; c0 < assign.size(); c0 += chunk) {
    const int nb = (int)std::min<size_t>(chunk, assign.size() - c0);
    for (int k = 0; k < nb; k++) moduli(assign[c0 + k], N, Emod.data() + V * k);
    check(m_batchCtx, femh_homogenize_many_modulus(m_batchCtx, nb, Emod.data(), C.data(), st.data()),
          "femh_homogenize_many_modulus");
    for (int k = 0; k < nb; k++) {
      for (int i = 0; i < 6; i++)
        for (int j = 0; j <= i; j++)
          oTensorValues.push_back((float)(m_referenceTriangle && j < i ? 0.0 : C[36 * k + 6 * i + j]));
      oParameters.push_back((float)computeMaterialDensity(assign[c0 + k], dens));
      if (iType == Cubic) {
        for (int q = 0; q < 36; q++) m_C.a[q] = C[36 * k + q];
        double E, nuc, G;
        cubicParameters(m_C, E, nuc, G);
        oParameters.push_back((float)E);
        oParameters.push_back((float)nuc);
        oParameters.push_back((float)G);
      }
    }
  }
}

void NumericalCoarsening::remember(int N[3], int rep, int sub, const std::vector<BaseMaterial> &mats) {
  m_key[0] = N[0];
  m_key[1] = N[1];
  m_key[2] = N[2];
  m_key[3] = rep;
  m_key[4] = sub;
  m_keyMats = mats;
}

void NumericalCoarsening::ensure(int N[3], int rep, int sub, const std::vector<BaseMaterial> &mats) {
  bool same = m_ctx && m_key[0] == N[0] && m_key[1] == N[1] && m_key[2] == N[2] && m_key[3] == rep &&
              m_key[4] == sub && m_keyMats.size() == mats.size();
  for (size_t i = 0; same && i < mats.size(); i++)
    same = m_keyMats[i].mu == mats[i].mu && m_keyMats[i].lambda == mats[i].lambda;
  if (!same) init(N, rep, sub, mats);
}

void NumericalCoarsening::toCoarseStrain(std::vector<double> &tv) {
  if (!m_coarseStrain || m_n[0] != m_n[1] || m_n[1] != m_n[2]) return;
  check(m_ctx, femh_coarse_strain_tensor(m_ctx, m_C.a.data()), "femh_coarse_strain_tensor");
  if (tv.size() < 21) return;
  size_t at = tv.size() - 21;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j <= i; j++) tv[at++] = m_referenceTriangle && j < i ? 0.0 : m_C(i, j);
}

Mat6 NumericalCoarsening::computeCoarsenedElasticityTensor(const std::vector<int> &iMaterials, int N[3], int rep,
                                                          int sub, const std::vector<BaseMaterial> &mats) {
  ensure(N, rep, sub, mats);
  std::vector<double> tv, pv;
  computeCoarsenedElasticityTensorAndParameters3D(iMaterials, N, {0.0, 1.0}, General, tv, pv);
  toCoarseStrain(tv);
  return m_C;
}

void NumericalCoarsening::computeCoarsenedElasticityTensorAndParameters(
    const std::vector<int> &iMaterials, int N[3], int rep, int sub, const std::vector<BaseMaterial> &mats,
    StructureType iType, std::vector<double> &ioTensorValues, std::vector<double> &ioParameters) {
  ensure(N, rep, sub, mats);
  std::vector<double> pv;
  computeCoarsenedElasticityTensorAndParameters3D(iMaterials, N, {0.0, 1.0}, General, ioTensorValues, pv);
  toCoarseStrain(ioTensorValues);
  // compliance columns: strains under a unit uniaxial stress along x, y, z
  const Mat6 Sfull = inverse6(m_C);
  double S[3][3];  // S[a][b]: strain b under a unit stress along a
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) S[a][b] = Sfull(b, a);
  double ratio = 0;
  for (int m : iMaterials) ratio += m == 1 ? 1.0 : 0.0;
  ioParameters.push_back(iMaterials.empty() ? 0.0 : ratio / (double)iMaterials.size());
  const int naxis = iType == Cubic ? 1 : 3;
  for (int a = 0; a < naxis; a++) ioParameters.push_back(1.0 / S[a][a]);
  ioParameters.push_back(-S[0][1] / S[0][0]);
  if (iType != Cubic) {
    ioParameters.push_back(-S[0][2] / S[0][0]);
    ioParameters.push_back(-S[1][2] / S[1][1]);
  }
  for (int a = 0; a < naxis; a++) ioParameters.push_back(m_C(3 + a, 3 + a));
}

// ---- caller pipeline files ------------------------------------------------------
namespace cfgbin {
namespace {
template <class T>
bool read_vec(std::ifstream &f, std::vector<T> &out) {
  int32_t n = 0;
  if (!f.read(reinterpret_cast<char *>(&n), sizeof(n)) || n < 0) return false;
  out.resize((size_t)n);
  if (n > 0 && !f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)(sizeof(T) * (size_t)n)))
    return false;
  return true;
}
template <class T>
bool write_vec(std::ofstream &f, const std::vector<T> &v) {
  const int32_t n = (int32_t)v.size();
  f.write(reinterpret_cast<const char *>(&n), sizeof(n));
  if (n > 0) f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(sizeof(T) * v.size()));
  return (bool)f;
}
template <class T>
bool read1(const std::string &file, std::vector<T> &out) {
  std::ifstream f(file, std::ios::binary);
  out.clear();
  return f.is_open() && read_vec(f, out);
}
template <class T>
bool read2(const std::string &file, std::vector<std::vector<T>> &out) {
  std::ifstream f(file, std::ios::binary);
  out.clear();
  int32_t nvec = 0;
  if (!f.is_open() || !f.read(reinterpret_cast<char *>(&nvec), sizeof(nvec)) || nvec < 0) return false;
  out.resize((size_t)nvec);
  for (auto &v : out)
    if (!read_vec(f, v)) return false;
  return true;
}
template <class T>
bool write1(const std::string &file, const std::vector<T> &v) {
  std::ofstream f(file, std::ios::binary | std::ios::trunc);
  return f.is_open() && write_vec(f, v);
}
template <class T>
bool write2(const std::string &file, const std::vector<std::vector<T>> &v) {
  std::ofstream f(file, std::ios::binary | std::ios::trunc);
  if (!f.is_open()) return false;
  const int32_t nvec = (int32_t)v.size();
  f.write(reinterpret_cast<const char *>(&nvec), sizeof(nvec));
  for (const auto &x : v)
    if (!write_vec(f, x)) return false;
  return (bool)f;
}
}  // namespace
bool readBinary(const std::string &file, std::vector<float> &out) { return read1(file, out); }
bool readBinary(const std::string &file, std::vector<int> &out) { return read1(file, out); }
bool readBinary(const std::string &file, std::vector<std::vector<int>> &out) { return read2(file, out); }
bool readBinary(const std::string &file, std::vector<std::vector<float>> &out) { return read2(file, out); }
bool writeBinary(const std::string &file, const std::vector<float> &v) { return write1(file, v); }
bool writeBinary(const std::string &file, const std::vector<int> &v) { return write1(file, v); }
bool writeBinary(const std::string &file, const std::vector<std::vector<int>> &v) { return write2(file, v); }
bool writeBinary(const std::string &file, const std::vector<std::vector<float>> &v) { return write2(file, v); }
}  // namespace cfgbin

int runTensorHomogenization(const std::string &dir, int level, int nbBlockRep, int nbSubdiv,
                            const std::vector<BaseMaterial> &mats, const std::vector<double> &dens,
                            bool cubicOnly, const femh_opts *solverOptions, int batch, int streams) {
  const std::string root = dir + (dir.empty() || dir.back() == '/' ? "" : "/") + "level" + std::to_string(level) + "_";
  std::vector<std::vector<int>> assign, baseMat;
  std::vector<float> params;
  if (!cfgbin::readBinary(root + "params.bin", params) || !cfgbin::readBinary(root + "baseMat.bin", baseMat) ||
      !cfgbin::readBinary(root + "matAssignments.bin", assign))
    throw HostError(FEMH_BAD_ARG, "cannot read " + root + "{params,baseMat,matAssignments}.bin");
  const NumericalCoarseningUsingMultigrid::StructureType type =
      cubicOnly ? NumericalCoarseningUsingMultigrid::Cubic : NumericalCoarseningUsingMultigrid::General;
  NumericalCoarseningUsingMultigrid nc;
  if (solverOptions) nc.setSolverOptions(*solverOptions);
  int n[3] = {level, level, level};
  nc.init(n, nbBlockRep, nbSubdiv, mats, type);
  std::vector<float> tensors, newParams;
  nc.computeCoarsenedElasticityTensorAndParameters3D(assign, n, dens, type, tensors, newParams, batch, streams);
  if (!cfgbin::writeBinary(root + "params.bin", newParams) || !cfgbin::writeBinary(root + "baseMat.bin", baseMat) ||
      !cfgbin::writeBinary(root + "matAssignments.bin", assign) ||
      !cfgbin::writeBinary(root + "elasticityTensors.bin", tensors))
    throw HostError(FEMH_BAD_ARG, "cannot write " + root + "*.bin");
  return (int)assign.size();
}

// ---- ThermalAnalysis (ThermalAnalysis.cpp:210-274, :388-505) -------------------
ThermalAnalysis::ThermalAnalysis() { femh_default_opts(&m_opts); }

bool ThermalAnalysis::run3D(int N[3], const std::vector<BaseMaterial> &mats,
                            const std::vector<std::vector<int>> &assign, const std::vector<float> &,
                            const std::vector<float> &, int rep, int sub, std::vector<float> &oCoef, int batch) {
  if (mats.size() < 2) throw HostError(FEMH_BAD_ARG, "need two base materials (void, solid)");
  if (batch < 1 || batch > 16) throw HostError(FEMH_BAD_ARG, "batch must be 1..16");
  oCoef.clear();
  m_alpha.clear();
  if (assign.empty()) return true;
  const int n[3] = {N[0] * rep * sub, N[1] * rep * sub, N[2] * rep * sub};
  double E1, nu1;
  mats[1].youngPoisson(E1, nu1);
  femh_opts o = m_opts;
  o.nx = n[0];
  o.ny = n[1];
  o.nz = n[2];
  o.E0 = E1;
  o.nu = nu1;
  o.interp = FEMH_INTERP_MODULUS;
  o.batch = batch;
  femh_ctx *ctx = nullptr;
  const int st = femh_create(&o, &ctx);
  if (st != FEMH_OK) {
    if (ctx) femh_destroy(ctx);
    fail(nullptr, st, "femh_create");
  }
  const size_t V = (size_t)n[0] * n[1] * n[2];
  std::vector<double> Emod(V * batch), alpha(V * batch), C(36 * (size_t)batch), aeff(6 * (size_t)batch);
  std::vector<int> cells;
  try {
    for (size_t c0 = 0; c0 < assign.size(); c0 += batch) {
      const int nb = (int)std::min<size_t>(batch, assign.size() - c0);
      for (int k = 0; k < nb; k++) {
        repMaterialAssignment(N[0], N[1], N[2], assign[c0 + k], rep, rep, rep, sub, cells);
        if (cells.size() != V) throw HostError(FEMH_BAD_ARG, "assignment does not match the grid");
        for (size_t e = 0; e < V; e++) {
          const bool solid = cells[e] == 1;
          Emod[V * k + e] = E1 * (solid ? 1.0 : m_densityMin);
          alpha[V * k + e] = solid ? 1.0 : 0.0;
        }
      }
      check(ctx, femh_homogenize_many_modulus(ctx, nb, Emod.data(), C.data(), nullptr), "femh_homogenize_many_modulus");
      check(ctx, femh_thermal_expansion(ctx, alpha.data(), nullptr, aeff.data()), "femh_thermal_expansion");
      for (int k = 0; k < nb; k++) {
        for (int q = 0; q < 6; q++) m_alpha.push_back(aeff[6 * k + q]);
        oCoef.push_back((float)((aeff[6 * k] + aeff[6 * k + 1] + aeff[6 * k + 2]) / 3.0));
      }
    }
  } catch (...) {
    femh_destroy(ctx);
    throw;
  }
  femh_destroy(ctx);
  return true;
}

int runThermalAnalysis(const std::string &dir, int level, int nbBlockRep, int nbSubdiv,
                       const std::vector<BaseMaterial> &mats, const femh_opts *solverOptions, int batch) {
  const std::string root = dir + (dir.empty() || dir.back() == '/' ? "" : "/") + "level" + std::to_string(level) + "_";
  std::vector<std::vector<int>> assign;
  if (!cfgbin::readBinary(root + "matAssignments.bin", assign))
    throw HostError(FEMH_BAD_ARG, "cannot read " + root + "matAssignments.bin");
  ThermalAnalysis ta;
  if (solverOptions) ta.setSolverOptions(*solverOptions);
  int n[3] = {level, level, level};
  std::vector<float> coef;
  ta.run3D(n, mats, assign, {}, {}, nbBlockRep, nbSubdiv, coef, batch);
  if (!cfgbin::writeBinary(root + "thermalExpansionCoefficients.bin", coef))
    throw HostError(FEMH_BAD_ARG, "cannot write " + root + "thermalExpansionCoefficients.bin");
  return (int)assign.size();
}

// ---- MaterialOptimizer (MaterialOptimizer.cpp:530-592, 621-652) ---------------
Mat6 MaterialOptimizer::computeTargetStrains(int idim, float E, float nu, float Gs) {
  if (idim != 3) throw HostError(FEMH_BAD_ARG, "only the 3D target is provided");
  Mat6 S;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) S(i, j) = (i == j ? 1.0 : -(double)nu) / (double)E;
  for (int i = 3; i < 6; i++) S(i, i) = 1.0 / (double)Gs;
  return S;
}

int MaterialOptimizer::lineSearch(RealFun *fun, VectorXd &x0, const VectorXd &dir, double &h) {
  const double f0 = fun->f();
  double f1 = f0;
  while (h > 1e-8) {
    VectorXd x(x0.size());
    for (size_t i = 0; i < x.size(); i++) x[i] = x0[i] - h * dir[i];
    clampVector(x, fun->lowerBounds, fun->upperBounds);
    fun->setParam(x);
    f1 = fun->f();
    if (f1 < f0) {
      x0 = x;
      h *= 2;
      break;
    }
    h /= 2;
  }
  return std::fabs(f0 - f1) < 1e-8 ? -1 : 0;
}

void MaterialOptimizer::gradientDescent(FEM3DFun *fun, VectorXd &x0, int nSteps, std::vector<VectorXd> &oParameters) {
  const double maxStep = 0.5;
  VectorXd x = x0;
  double h = 1;
  fun->setParam(x);
  for (int ii = 0; ii < nSteps; ii++) {
    const VectorXd grad = fun->df();
    const double norm = infNorm(grad);
    if (!(norm > 0)) break;
    if (ii == 0) h = maxStep / norm;
    if (lineSearch(fun, x, grad, h) < 0) break;
    x0 = x;
    oParameters.push_back(fun->distribution);
  }
}

bool MaterialOptimizer::run3D(int N[3], const std::vector<BaseMaterial> &mats, const std::vector<int> &assign,
                              const std::vector<float> &current, const std::vector<float> &target,
                              std::vector<std::vector<int>> &out, int nSteps) {
  (void)current;
  if (target.size() < 4) throw HostError(FEMH_BAD_ARG, "target parameters are (density, E, nu, G)");
  const int nx = N[0], ny = N[1], nz = N[2];
  if ((int)assign.size() != nx * ny * nz) throw HostError(FEMH_BAD_ARG, "assignment size");
  FEM3DFun fem;
  if (m_haveOpts) fem.solverOpts = m_opts;
  double E, nu;
  mats.at(1).youngPoisson(E, nu);
  fem.nu = nu;
  PiecewiseConstant3D plain;
  PiecewiseConstantSym3D sym;
  PiecewiseConstantCubic3D cubic;
  RealField *field = &plain;
  PiecewiseConstant3D *pc = &plain;
  if (m_structureType == Cubic) pc = &cubic;
  if (m_structureType == Orthotropic) pc = &sym;
  field = pc;
  // symmetric fields carry one octant (nx/2 per axis); the plain field the full grid
  const bool half = m_structureType != General;
  const int px = half ? nx / 2 : nx, py = half ? ny / 2 : ny, pz = half ? nz / 2 : nz;
  pc->allocate(px, py, pz);
  fem.field = field;
  fem.gridSize = {nx, ny, nz};
  fem.keepDisplacements = false;  // the loop reads G only
  fem.lowerBounds.assign(field->param.size(), 1e-3);
  fem.upperBounds.assign(field->param.size(), 1.0);
  // start from the shrunk 0/1 assignment (the octant for symmetric fields)
  VectorXd x0(field->param.size());
  for (int i = 0; i < px; i++)
    for (int j = 0; j < py; j++)
      for (int k = 0; k < pz; k++) {
        const int m = assign[((size_t)i * ny + j) * nz + k] == 0 ? 0 : 1;
        x0[((size_t)i * py + j) * pz + k] = 0.5 + (m - 0.5) * 0.3;
      }
  fem.initArrays();
  fem.init(x0);
  fem.setParam(x0);
  // target: the cubic stiffness of (E, nu, G), matched on C00, C01, C33
  const Mat6 Ct = inverse6(computeTargetStrains(3, target[1], target[2], target[3]));
  fem.G0 = fem.G;
  fem.G0(0, 0) = Ct(0, 0);
  fem.G0(0, 1) = Ct(0, 1);
  fem.G0(3, 3) = Ct(3, 3);
  fem.wG(0, 0) = fem.wG(0, 1) = fem.wG(3, 3) = 1.0;
  fem.m0 = target[0];
  fem.mw = 0.1 * fem.G(0, 0) / std::max(fem.density, 1e-12);
  std::vector<VectorXd> dists;
  gradientDescent(&fem, x0, nSteps, dists);
  for (const VectorXd &d : dists) {
    std::vector<int> a(d.size());
    for (size_t e = 0; e < d.size(); e++) a[e] = d[e] >= 0.5 ? 1 : 0;
    out.push_back(a);
  }
  return !dists.empty();
}

}  // namespace femh
