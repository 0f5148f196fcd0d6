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
  std::printf("%s\n", fails ? "FAILED" : "ALL PASSED");
  return fails ? 1 : 0;
}
