// test_host.cpp -- exercises the C++ host facade (fem_amd/host/femh_host.hpp)
// the way reference callers use its classes; needs a GPU.  Prints one line
// per check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../fem_amd/host/femh_host.hpp"

using namespace femh;

static int fails = 0;
static void expect(bool ok, const char *what, double v) {
  std::printf("%s %s (%.3e)\n", ok ? "PASS" : "FAIL", what, v);
  std::fflush(stdout);
  if (!ok) fails++;
}

// isotropic stiffness (Voigt, engineering shear)
static Mat6 c_iso(double E, double nu) {
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
  Mat6 C;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C(i, j) = lam + (i == j ? 2 * mu : 0.0);
  for (int i = 3; i < 6; i++) C(i, i) = mu;
  return C;
}

// x = C^-1 b by Gaussian elimination with partial pivoting
static void solve6(const Mat6 &C, const double *b, double *x) {
  double M[6][7];
  for (int i = 0; i < 6; i++) {
    for (int j = 0; j < 6; j++) M[i][j] = C(i, j);
    M[i][6] = b[i];
  }
  for (int c = 0; c < 6; c++) {
    int p = c;
    for (int r = c + 1; r < 6; r++)
      if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
    for (int j = 0; j < 7; j++) std::swap(M[c][j], M[p][j]);
    for (int r = c + 1; r < 6; r++) {
      const double f = M[r][c] / M[c][c];
      for (int j = c; j < 7; j++) M[r][j] -= f * M[c][j];
    }
  }
  for (int i = 5; i >= 0; i--) {
    double v = M[i][6];
    for (int j = i + 1; j < 6; j++) v -= M[i][j] * x[j];
    x[i] = v / M[i][i];
  }
}

int main() {
  // base materials as the sample generator builds them: mu = 100,
  // lambda = 1000, the void with E / 1000 (SamplesGeneratorImpl.cpp:4864-4893)
  std::vector<BaseMaterial> mats = {{0.1, 1.0}, {100.0, 1000.0}};
  double E1, nu1;
  mats[1].youngPoisson(E1, nu1);

  {  // all-solid cell: C_H is the base material's isotropic tensor
    NumericalCoarseningUsingMultigrid nc;
    int N[3] = {4, 4, 4};
    nc.init(N, 1, 2, mats, NumericalCoarseningUsingMultigrid::Cubic);
    std::vector<int> m(64, 1);
    std::vector<double> tv, par;
    nc.computeCoarsenedElasticityTensorAndParameters3D(m, N, {0.0, 1.0}, NumericalCoarseningUsingMultigrid::Cubic,
                                                       tv, par);
    const Mat6 Ci = c_iso(E1, nu1);
    double err = 0;
    int p = 0;
    for (int i = 0; i < 6; i++)
      for (int j = 0; j <= i; j++) err = std::fmax(err, std::fabs(tv[p++] - Ci(i, j)) / Ci(0, 0));
    expect(tv.size() == 21 && err < 1e-12, "coarsening: homogeneous cell gives C_iso", err);
    expect(par.size() == 4 && std::fabs(par[1] - E1) < 1e-9 * E1 && std::fabs(par[2] - nu1) < 1e-9,
           "coarsening: cubic parameters recover (E, nu)", std::fabs(par[1] - E1) / E1);
  }
  {  // NumericalCoarsening's tensor-only and force-path parameter calls on a solid cell
    NumericalCoarsening ncf;
    int N[3] = {4, 4, 4};
    std::vector<int> m(64, 1);
    const Mat6 C = ncf.computeCoarsenedElasticityTensor(m, N, 1, 2, mats);
    const Mat6 Ci = c_iso(E1, nu1);
    double err = 0;
    for (int q = 0; q < 36; q++) err = std::fmax(err, std::fabs(C.a[q] - Ci.a[q]) / Ci(0, 0));
    expect(err < 1e-12, "coarsening: computeCoarsenedElasticityTensor of a solid cell is C_iso", err);
    std::vector<double> tv, pv;
    ncf.computeCoarsenedElasticityTensorAndParameters(m, N, 1, 2, mats, NumericalCoarsening::General, tv, pv);
    const double mu1 = mats[1].mu;
    // ratio, E x3, nu_xy, nu_xz, nu_yz, G x3
    double perr = std::fabs(pv[0] - 1.0);
    for (int a = 0; a < 3; a++) perr = std::fmax(perr, std::fabs(pv[1 + a] - E1) / E1);
    for (int a = 0; a < 3; a++) perr = std::fmax(perr, std::fabs(pv[4 + a] - nu1));
    for (int a = 0; a < 3; a++) perr = std::fmax(perr, std::fabs(pv[7 + a] - mu1) / mu1);
    expect(tv.size() == 21 && pv.size() == 10 && perr < 1e-10,
           "coarsening: force-path parameters (ratio, E, nu, G) of a solid cell", perr);
  }
  {  // a cubic-symmetric pattern: symmetric, cubic invariants, reference triangle quirk
    NumericalCoarseningUsingMultigrid nc;
    int N[3] = {4, 4, 4};
    nc.init(N, 2, 1, mats, NumericalCoarseningUsingMultigrid::Cubic);
    std::vector<int> m(64, 0);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++)
        for (int k = 0; k < 4; k++) {
          const int c = (i == 0) + (j == 0) + (k == 0);
          m[(i * 4 + j) * 4 + k] = c >= 2 ? 1 : 0;  // a frame of edges
        }
    std::vector<double> tv, par;
    nc.computeCoarsenedElasticityTensorAndParameters3D(m, N, {0.0, 1.0}, NumericalCoarseningUsingMultigrid::Cubic,
                                                       tv, par);
    const Mat6 &C = nc.lastTensor();
    const double c0 = C(0, 0);
    const double asym = std::fmax(std::fabs(C(1, 1) - c0), std::fabs(C(2, 2) - c0)) / c0 +
                        std::fabs(C(0, 1) - C(1, 2)) / c0 + std::fabs(C(3, 3) - C(5, 5)) / c0;
    // 10 of the 64 cells have at least two zero coordinates
    expect(asym < 1e-8 && std::fabs(par[0] - 10.0 / 64.0) < 1e-15, "coarsening: frame lattice is cubic",
           asym);
    // Cubic took the two-solve path on this swap-invariant cell; General
    // solves all six cases
    NumericalCoarseningUsingMultigrid ng;
    ng.init(N, 2, 1, mats, NumericalCoarseningUsingMultigrid::General);
    std::vector<double> tg, pg;
    ng.computeCoarsenedElasticityTensorAndParameters3D(m, N, {0.0, 1.0}, NumericalCoarseningUsingMultigrid::General,
                                                       tg, pg);
    double dgc = 0;
    for (int q = 0; q < 21; q++) dgc = std::fmax(dgc, std::fabs(tg[q] - tv[q]) / c0);
    expect(dgc < 1e-8, "coarsening: Cubic (two solves + axis swaps) = General (six solves)", dgc);
    nc.setEmitReferenceTriangle(true);
    std::vector<double> tq, pq;
    nc.computeCoarsenedElasticityTensorAndParameters3D(m, N, {0.0, 1.0}, NumericalCoarseningUsingMultigrid::Cubic,
                                                       tq, pq);
    expect(tq[1] == 0.0 && tq[0] == tv[0], "coarsening: reference triangle output is diagonal", tq[1]);
  }
  {  // FEM3DFun: adjoint gradient against central differences through the facade
    FEM3DFun fem;
    PiecewiseConstant3D field;
    field.allocate(4, 4, 4);
    fem.field = &field;
    fem.gridSize = {8, 8, 8};
    fem.solverOpts.tol = 1e-13;
    fem.solverOpts.max_iter = 20000;
    fem.initArrays();
    fem.lowerBounds.assign(64, 1e-3);
    fem.upperBounds.assign(64, 1.0);
    std::vector<double> x0(64);
    for (int i = 0; i < 64; i++) x0[i] = 0.3 + 0.6 * std::fabs(std::sin(1.7 * i + 0.3));
    fem.init(x0);
    fem.setParam(x0);
    {  // the reference's own energy loop over u and Ke0 (fast_fem/FEM3DFun.cpp:160-177)
      const int n = 8, N1 = n + 1;
      double gerr = 0;
      Mat6 Gr;
      for (int c = 0; c < n * n * n; c++) {
        const int i = c / (n * n), j = (c / n) % n, k = c % n;
        const double dist = fem.distribution[c], coef = dist / (3 - 2 * dist);
        double ue[6][24];
        for (int lc = 0; lc < 6; lc++)
          for (int v = 0; v < 8; v++)
            for (int d = 0; d < 3; d++) {
              const int vi = i + ((v >> 2) & 1), vj = j + ((v >> 1) & 1), vk = k + (v & 1);
              ue[lc][3 * v + d] = fem.u[lc][3 * ((vi * N1 + vj) * N1 + vk) + d];
            }
        for (int a = 0; a < 6; a++)
          for (int b = a; b < 6; b++) {
            double e = 0;
            for (int r = 0; r < 24; r++)
              for (int q = 0; q < 24; q++) e += ue[a][r] * fem.Ke0[0][24 * r + q] * ue[b][q];
            Gr(a, b) += coef * e;
          }
      }
      for (int q = 0; q < 36; q++) gerr = std::fmax(gerr, std::fabs(Gr.a[q] - fem.G.a[q]) / fem.G(0, 0));
      expect(fem.u.size() == 6 && fem.u[0].size() == 3u * N1 * N1 * N1 && gerr < 1e-10,
             "FEM3DFun: u and Ke0 in the reference layout reproduce G", gerr);
    }
    fem.G0 = fem.G;
    fem.G0(0, 0) *= 1.3;
    fem.G0(0, 1) *= 0.5;
    fem.G0(3, 3) *= 1.1;
    fem.wG(0, 0) = fem.wG(0, 1) = fem.wG(3, 3) = 1.0;
    fem.m0 = 0.4;
    fem.mw = 0.1;
    const double rel = check_df(&fem, x0, 1e-6);
    expect(rel < 1e-5, "FEM3DFun: df matches central differences (h = 1e-6)", rel);
    fem.setParam(x0);
    const double f0 = fem.f();
    std::vector<double> x = x0;
    gradientDescent(&fem, x, 5, 0.1, 1, nullptr);
    fem.setParam(x);
    const double f1 = fem.f();
    expect(f1 < f0, "gradientDescent decreases the objective", f0 - f1);
  }
  {  // MaterialOptimizer on a cubic field
    MaterialOptimizer opt;
    opt.setStructureType(MaterialOptimizer::Cubic);
    int N[3] = {8, 8, 8};
    std::vector<int> a(512);
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++)
        for (int k = 0; k < 8; k++) a[(i * 8 + j) * 8 + k] = ((i < 4) + (j < 4) + (k < 4)) >= 2;
    std::vector<std::vector<int>> out;
    const bool ok = opt.run3D(N, mats, a, {0.5f, 1.0f, 0.3f, 0.3f}, {0.4f, 50.0f, 0.2f, 20.0f}, out, 3);
    expect(ok && !out.empty() && out[0].size() == 512, "MaterialOptimizer::run3D returns assignments",
           (double)out.size());
  }
  {  // errors surface as HostError with the library's status
    FEM3DFun fem;
    PiecewiseConstant3D field;
    field.allocate(1, 1, 1);
    fem.field = &field;
    fem.gridSize = {1, 4, 4};
    bool threw = false;
    try {
      fem.setParam({0.5});
    } catch (const HostError &e) {
      threw = e.status == FEMH_BAD_ARG;
    }
    expect(threw, "bad grid raises HostError(BAD_ARG)", 0.0);
  }
  {  // the tensor-homogenization stage as a batch job over the caller's files
    const char *tmp = std::getenv("TMPDIR");
    const std::string dir = std::string(tmp && *tmp ? tmp : "/tmp");
    const int level = 4, nsamp = 6;
    std::vector<std::vector<int>> assign(nsamp, std::vector<int>(level * level * level, 0));
    unsigned seed = 12345u;
    for (auto &a : assign)
      for (auto &m : a) {
        seed = seed * 1664525u + 1013904223u;
        m = (seed >> 16) % 3 != 0 ? 1 : 0;
      }
    const std::string root = dir + "/level4_";
    bool ok = cfgbin::writeBinary(root + "matAssignments.bin", assign) &&
              cfgbin::writeBinary(root + "baseMat.bin", std::vector<std::vector<int>>{{0}, {1}}) &&
              cfgbin::writeBinary(root + "params.bin", std::vector<float>(nsamp * 4, 0.0f));
    femh_opts o;
    femh_default_opts(&o);
    o.tol = 1e-12;
    const int n = ok ? runTensorHomogenization(dir, level, 1, 2, mats, {0.0, 1.0}, true, &o, 4) : 0;
    std::vector<float> tens, par;
    std::vector<std::vector<int>> back;
    ok = ok && n == nsamp && cfgbin::readBinary(root + "elasticityTensors.bin", tens) &&
         cfgbin::readBinary(root + "params.bin", par) && cfgbin::readBinary(root + "matAssignments.bin", back);
    expect(ok && tens.size() == 21u * nsamp && par.size() == 4u * nsamp && back == assign,
           "runTensorHomogenization: files read and written in the caller's layout", (double)tens.size());
    // against one-cell calls of the same facade
    NumericalCoarseningUsingMultigrid nc;
    nc.setSolverOptions(o);
    int N[3] = {level, level, level};
    nc.init(N, 1, 2, mats, NumericalCoarseningUsingMultigrid::Cubic);
    double err = 0, perr = 0;
    for (int s = 0; s < nsamp && ok; s++) {
      std::vector<double> tv, pv;
      nc.computeCoarsenedElasticityTensorAndParameters3D(assign[s], N, {0.0, 1.0},
                                                         NumericalCoarseningUsingMultigrid::Cubic, tv, pv);
      for (int q = 0; q < 21; q++) err = std::fmax(err, std::fabs(tens[21 * s + q] - tv[q]) / std::fabs(tv[0]));
      for (int q = 0; q < 4; q++) perr = std::fmax(perr, std::fabs(par[4 * s + q] - pv[q]) / (std::fabs(pv[q]) + 1e-30));
    }
    expect(ok && err < 1e-6, "runTensorHomogenization: batched tensors = one-cell tensors (float)", err);
    expect(ok && perr < 1e-5, "runTensorHomogenization: cubic parameters = one-cell parameters", perr);
    for (const char *f : {"matAssignments", "baseMat", "params", "elasticityTensors"})
      std::remove((root + f + ".bin").c_str());
  }
  {  // thermal expansion: a solid cell expands by 1; a two-phase cell obeys Levin's relation
    femh_opts o;
    femh_default_opts(&o);
    o.tol = 1e-12;
    int N[3] = {4, 4, 4};
    std::vector<std::vector<int>> assign(2, std::vector<int>(64, 1));
    unsigned seed = 777u;
    for (auto &m : assign[1]) {
      seed = seed * 1664525u + 1013904223u;
      m = (seed >> 16) % 2;
    }
    ThermalAnalysis ta;
    ta.setSolverOptions(o);
    std::vector<float> coef;
    ta.run3D(N, mats, assign, {}, {}, 1, 2, coef, 2);
    const std::vector<double> &a = ta.lastExpansions();
    expect(coef.size() == 2 && a.size() == 12 && std::fabs(a[0] - 1.0) < 1e-10 && std::fabs(a[3]) < 1e-10,
           "thermal: a solid cell expands uniformly", std::fabs(a[0] - 1.0));
    NumericalCoarseningUsingMultigrid nc;
    nc.setSolverOptions(o);
    nc.init(N, 1, 2, mats, NumericalCoarseningUsingMultigrid::General);
    std::vector<double> tv, pv;
    nc.computeCoarsenedElasticityTensorAndParameters3D(assign[1], N, {0.0, 1.0},
                                                       NumericalCoarseningUsingMultigrid::General, tv, pv);
    // Levin: alpha = phi I + (S I - <1/3K> I) / (1/3K1 - 1/3K2), K2 = 1e-4 K1
    const double ones[6] = {1, 1, 1, 0, 0, 0};
    double SI6[6];
    solve6(nc.lastTensor(), ones, SI6);
    const double k1 = E1 / (1 - 2 * nu1), k2 = 1e-4 * k1;  // 3K of each phase
    double phi = 0;
    for (int m : assign[1]) phi += m;
    phi /= 64.0;
    const double mc = phi / k1 + (1 - phi) / k2;
    double lerr = 0;
    for (int i = 0; i < 6; i++) {
      const double SI = SI6[i], Ii = i < 3 ? 1.0 : 0.0;
      const double lev = phi * Ii + (SI - mc * Ii) / (1 / k1 - 1 / k2);
      lerr = std::fmax(lerr, std::fabs(a[6 + i] - lev));
    }
    expect(lerr < 1e-7, "thermal: two-phase cell obeys Levin's relation", lerr);
    expect(std::fabs(coef[1] - (a[6] + a[7] + a[8]) / 3.0) < 1e-6, "thermal: coefficient is the mean normal expansion",
           coef[1]);
    // the stage as a batch job over the caller's files
    const char *tmp = std::getenv("TMPDIR");
    const std::string dir = std::string(tmp && *tmp ? tmp : "/tmp");
    const std::string root = dir + "/level4_";
    bool ok = cfgbin::writeBinary(root + "matAssignments.bin", assign);
    const int cnt = ok ? runThermalAnalysis(dir, 4, 1, 2, mats, &o, 2) : 0;
    std::vector<float> back;
    ok = ok && cnt == 2 && cfgbin::readBinary(root + "thermalExpansionCoefficients.bin", back);
    expect(ok && back == coef, "runThermalAnalysis: coefficients written in the caller's layout", (double)back.size());
    for (const char *f : {"matAssignments", "thermalExpansionCoefficients"}) std::remove((root + f + ".bin").c_str());
  }
  std::printf("%s\n", fails ? "FAILED" : "ALL PASSED");
  return fails ? 1 : 0;
}
