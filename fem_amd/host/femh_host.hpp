// femh_host.hpp -- C++ host facade over the femh C ABI (include/femh.h) that
// keeps desaic/fem's class names and call shapes for the homogenization path,
// so that callers written against the reference switch by changing includes
// and the vector type (std::vector<double> instead of Eigen::VectorXd):
//
//   RealFun                          MaterialOpt/include/RealFun.hpp:13-52
//   RealField, PiecewiseConstant3D,  MaterialOpt/src/PiecewiseConstant3D.cpp,
//   PiecewiseConstantSym3D,          PiecewiseConstantSym3D.cpp,
//   PiecewiseConstantCubic3D         PiecewiseConstantCubic3D.cpp
//   FEM3DFun : RealFun               fast_fem/FEM3DFun.hpp:13-92,
//                                    fast_fem/FEM3DFun.cpp:128-268
//   lineSearch, gradientDescent,     fast_fem/run3d.cpp:238-300
//   check_df
//   NumericalCoarseningUsingMultigrid SamplesGenerator/NumericalCoarseningUsingMultigrid.h:11-48,
//                                    NumericalCoarseningUsingMultiGrid.cpp:134-340
//   NumericalCoarsening              SamplesGenerator/NumericalCoarsening.h:13-86 (3D entry)
//   MaterialOptimizer                SamplesGenerator/MaterialOptimizer.h:13-60,
//                                    MaterialOptimizer.cpp:530-592
//
// Differences from the reference, on purpose:
//   * errors raise femh::HostError (with the library's status and message)
//     instead of asserts / printf;
//   * the solve is the engine's strain-driven periodic homogenization (GPU
//     multigrid-PCG) for both coarsening classes;
//   * NumericalCoarseningUsingMultigrid writes the full lower triangle of the
//     symmetric C_H; the reference fills C(i, j) for j >= i and then emits
//     j <= i, i.e. only the diagonal (NumericalCoarseningUsingMultiGrid.cpp:
//     297-325).  setEmitReferenceTriangle(true) reproduces that output.
#pragma once
#include <array>
#include <ostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/femh.h"

namespace femh {

struct HostError : std::runtime_error {
  int status;
  HostError(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

using VectorXd = std::vector<double>;

// 6x6 row-major matrix (Voigt xx, yy, zz, yz, xz, xy)
struct Mat6 {
  std::array<double, 36> a{};
  double &operator()(int i, int j) { return a[6 * i + j]; }
  double operator()(int i, int j) const { return a[6 * i + j]; }
};

// One base material of the reference's MaterialQuad list, by its Lame
// parameters (MaterialQuad::e[0]->param = {mu, lambda}).
struct BaseMaterial {
  double mu, lambda;
  void youngPoisson(double &E, double &nu) const;  // cfgMaterialUtilities.cpp:2935-2944
};

// ---- optimisation interface ---------------------------------------------
class RealFun {
 public:
  VectorXd param;
  // may be empty: unbounded
  VectorXd lowerBounds, upperBounds;
  virtual void init(const VectorXd &x0) { param = x0; }
  virtual void setParam(const VectorXd &x0) { param = x0; }
  virtual double f() = 0;
  virtual VectorXd df();
  virtual void log(std::ostream &) {}
  virtual ~RealFun() {}
};

// material distribution over the unit cube from a parameter vector; df
// returns the (index, weight) entries of d f(x) / d param
class RealField {
 public:
  VectorXd param;
  virtual double f(const double x[3]) = 0;
  virtual std::vector<std::pair<int, double>> df(const double x[3]) = 0;
  void setParam(int i, double v) { param[i] = v; }
  virtual ~RealField() {}
};

class PiecewiseConstant3D : public RealField {
 public:
  int gridSize[3] = {0, 0, 0};
  void allocate(int nx, int ny, int nz);
  double f(const double x[3]) override;
  std::vector<std::pair<int, double>> df(const double x[3]) override;

 protected:
  virtual void gridIdx(const double x[3], int idx[3]) const;
};

// mirror-symmetric in each axis about the cell centre
class PiecewiseConstantSym3D : public PiecewiseConstant3D {
 protected:
  void gridIdx(const double x[3], int idx[3]) const override;
};

// mirror-symmetric and invariant under axis permutations
class PiecewiseConstantCubic3D : public PiecewiseConstant3D {
 protected:
  void gridIdx(const double x[3], int idx[3]) const override;
};

// Objective 0.5 sum wG (G - G0)^2 + 0.5 mw (density - m0)^2 over the
// homogenized tensor G of the distribution clamp(field(x), 1e-3, 1), with the
// fast_fem interpolation E = rho / (3 - 2 rho) of a unit base material.
class FEM3DFun : public RealFun {
 public:
  FEM3DFun();
  ~FEM3DFun() override;
  FEM3DFun(const FEM3DFun &) = delete;
  FEM3DFun &operator=(const FEM3DFun &) = delete;

  // G(i, j) for j >= i (the upper triangle the reference computes), 0 below
  Mat6 G, G0, wG;
  double density = 0, m0 = 0, mw = 0;
  std::vector<int> gridSize;  // voxels per axis
  double nu = 0.45;           // base material (E = 1), as fast_fem/run3d.cpp:156-157
  femh_opts solverOpts;       // knobs of the engine (grid/material fields are set here)
  RealField *field = nullptr;  // not owned
  VectorXd distribution;
  // total displacements of the 6 load cases after setParam, on the
  // (n+1)^3 nodes of the unit cell, 3 * ((i (n+1) + j) (n+1) + k) + d, in
  // cell coordinates X = dx (i, j, k) with dx = 1 / n: u = dx (eps X / dx +
  // chi(i mod n, j mod n, k mod n)), the layout of copyPeriodicU
  // (fast_fem/FEM3DFun.cpp:31-55, :157-159).  Filled when keepDisplacements.
  std::vector<VectorXd> u;
  bool keepDisplacements = true;
  // base-material element stiffness for cells of size dx (E = 1): Ke0[0],
  // 24 x 24 row-major, corner order z fastest, 3 dofs per corner
  std::vector<VectorXd> Ke0;

  void initArrays();
  void init(const VectorXd &x0) override;
  void setParam(const VectorXd &x0) override;
  double f() override;
  VectorXd df() override;
  void log(std::ostream &out) override;
  // periodic fluctuation fields of the last setParam: [6][3][V]
  VectorXd fields();
  const femh_stats &lastStats() const { return stats_; }

 private:
  femh_ctx *ctx_ = nullptr;
  femh_stats stats_{};
  void ensureContext();
};

int lineSearch(RealFun *fun, VectorXd &x0, const VectorXd &dir, double &h);
void gradientDescent(RealFun *fun, VectorXd &x0, int nSteps, double maxStep, int logInterval,
                     std::ostream *logfile = nullptr);
// max |analytic - central difference| / max |analytic|
double check_df(RealFun *fun, const VectorXd &x0, double h);
double infNorm(const VectorXd &a);
void clampVector(VectorXd &x, const VectorXd &lb, const VectorXd &ub);

// ---- coarsening ------------------------------------------------------------
class NumericalCoarseningUsingMultigrid {
 public:
  enum StructureType { General, Cubic, Orthotropic };
  NumericalCoarseningUsingMultigrid();
  ~NumericalCoarseningUsingMultigrid();
  NumericalCoarseningUsingMultigrid(const NumericalCoarseningUsingMultigrid &) = delete;
  NumericalCoarseningUsingMultigrid &operator=(const NumericalCoarseningUsingMultigrid &) = delete;

  void setUseVariableDensity(bool v) { m_useVariableDensity = v; }
  void setEmitReferenceTriangle(bool v) { m_referenceTriangle = v; }
  void setSolverOptions(const femh_opts &o) { m_opts = o; }

  // Cubic on a cube grid selects the engine's cubic-structure mode, as the
  // reference's init selects its CUBIC solve (two load cases for cells
  // invariant under the axis swaps; the engine solves other cells in full)
  void init(int N[3], int iNbBlockRep, int iNbSubdiv, const std::vector<BaseMaterial> &iBaseMaterials,
            StructureType iType);
  // appends 21 values (row i, j <= i) to ioTensorValues and, for Cubic,
  // (density, E, nu, G) to ioParameters
  void computeCoarsenedElasticityTensorAndParameters3D(const std::vector<int> &iMaterials, int N[3],
                                                       const std::vector<double> &iBaseMaterialDensities,
                                                       StructureType iType,
                                                       std::vector<double> &ioTensorValues,
                                                       std::vector<double> &ioParameters);
  // the full tensor of the last call
  const Mat6 &lastTensor() const { return m_C; }

  // All assignments at once, `batch` cells per launch sequence, `streams`
  // batches in flight (extra engine contexts of that batch size): per assignment, in order, the same
  // 21 values and (Cubic) 4 parameters as the one-cell call, as float (the
  // caller pipeline's file type).  General appends the density only (the
  // reference's MG path has no General parameters: NumericalCoarseningUsing-
  // MultiGrid.cpp:342-345).
  void computeCoarsenedElasticityTensorAndParameters3D(const std::vector<std::vector<int>> &iMaterialAssignments,
                                                       int N[3], const std::vector<double> &iBaseMaterialDensities,
                                                       StructureType iType, std::vector<float> &oTensorValues,
                                                       std::vector<float> &oParameters, int batch = 16,
                                                       int streams = 3);

 protected:
  femh_ctx *m_ctx = nullptr;
  femh_opts m_opts;
  int m_n[3] = {0, 0, 0};
  int m_nbBlockRep = 1, m_nbSubdiv = 1;
  std::vector<BaseMaterial> m_baseMaterials;
  bool m_useVariableDensity = true;
  bool m_referenceTriangle = false;
  double m_densityMin = 1e-4;
  Mat6 m_C;
  femh_ctx *m_batchCtx = nullptr;
  int m_batch = 0, m_streams = 0;
  int m_structure = FEMH_STRUCT_GENERAL;  // engine structure chosen by init()
  void moduli(const std::vector<int> &iMaterials, int N[3], double *oE) const;
};

// Same entry point as the reference's PARDISO-based coarsening.  It runs the
// engine's periodic strain-driven solve; on a cube cell the tensor it returns
// is the force-driven path's coarse-strain form (femh_coarse_strain_tensor:
// what the reference's face-traction KKT solve and
// MeshUtilities::computeCoarsenedElasticityTensor give, without the 1e-5
// diagonal regularisation); on other cells, and with
// setCoarseStrainTensor(false), the energy form of the multigrid path.
class NumericalCoarsening : public NumericalCoarseningUsingMultigrid {
 public:
  void setCoarseStrainTensor(bool v) { m_coarseStrain = v; }
  void init(int N[3], int iNbBlockRep, int iNbSubdiv, const std::vector<BaseMaterial> &iBaseMaterials) {
    NumericalCoarseningUsingMultigrid::init(N, iNbBlockRep, iNbSubdiv, iBaseMaterials, General);
    remember(N, iNbBlockRep, iNbSubdiv, iBaseMaterials);
  }
  // C_H of one assignment (NumericalCoarsening.cpp:577-596); initialises for
  // this grid and these materials when they differ from the last call
  Mat6 computeCoarsenedElasticityTensor(const std::vector<int> &iMaterials, int N[3], int iNbBlockRep,
                                        int iNbSubdiv, const std::vector<BaseMaterial> &iBaseMaterials);
  // 21 values of C_H and the parameters of the reference's force-driven path
  // (:486-552), read from the compliance S = C_H^-1 (its columns are the
  // strains under a unit uniaxial stress): the material-1 ratio, Young's
  // moduli 1/S_ii (x only for Cubic), nu_xy = -S_10/S_00 (and, unless Cubic,
  // nu_xz = -S_20/S_00, nu_yz = -S_21/S_11), shear moduli C_33 (and C_44,
  // C_55 unless Cubic)
  void computeCoarsenedElasticityTensorAndParameters(const std::vector<int> &iMaterials, int N[3],
                                                     int iNbBlockRep, int iNbSubdiv,
                                                     const std::vector<BaseMaterial> &iBaseMaterials,
                                                     StructureType iType, std::vector<double> &ioTensorValues,
                                                     std::vector<double> &ioParameters);
  using NumericalCoarseningUsingMultigrid::computeCoarsenedElasticityTensorAndParameters3D;

 private:
  int m_key[5] = {0, 0, 0, 0, 0};
  std::vector<BaseMaterial> m_keyMats;
  bool m_coarseStrain = true;
  void remember(int N[3], int rep, int sub, const std::vector<BaseMaterial> &mats);
  void ensure(int N[3], int rep, int sub, const std::vector<BaseMaterial> &mats);
  // replace the energy tensor of the last solve (m_C and the 21 values last
  // appended to tv) by the coarse-strain form where it applies
  void toCoarseStrain(std::vector<double> &tv);
};

// replicate a coarse assignment rep times per axis with each cell split into
// subdiv^3 (cfgMaterialUtilities.cpp:920-960)
void repMaterialAssignment(int nx, int ny, int nz, const std::vector<int> &iMaterialCombIndices, int repX,
                           int repY, int repZ, int iNbSubdivisions, std::vector<int> &oMaterials);
double computeMaterialDensity(const std::vector<int> &iMaterialAssignment,
                              const std::vector<double> &iBaseMaterialDensities);
// cubic (E, nu, G) from C (NumericalCoarseningUsingMultiGrid.cpp:327-340)
void cubicParameters(const Mat6 &C, double &E, double &nu, double &G);

// ---- caller pipeline files -----------------------------------------------------
// The sample generator's binary layout (Utilities/include/cfgUtilities.h:
// 803-870), host byte order: a vector is an int32 count then the raw
// elements; a vector of vectors is an int32 count then each vector as above.
// read* return false on a missing or truncated file.
namespace cfgbin {
bool readBinary(const std::string &file, std::vector<float> &out);
bool readBinary(const std::string &file, std::vector<int> &out);
bool readBinary(const std::string &file, std::vector<std::vector<int>> &out);
bool readBinary(const std::string &file, std::vector<std::vector<float>> &out);
bool writeBinary(const std::string &file, const std::vector<float> &v);
bool writeBinary(const std::string &file, const std::vector<int> &v);
bool writeBinary(const std::string &file, const std::vector<std::vector<int>> &v);
bool writeBinary(const std::string &file, const std::vector<std::vector<float>> &v);
}  // namespace cfgbin

// SamplesGeneratorImpl::runTensorHomogenization (SamplesGeneratorImpl.cpp:
// 4613-4631, readFiles :3854-3923 binary branch, writeFiles :3826-3838) as a
// GPU batch job: reads dir/level{L}_{params,baseMat,matAssignments}.bin,
// homogenizes every assignment of the L^3 grid (replicated nbBlockRep times
// per axis, each cell split into nbSubdiv^3 voxels), and writes
// dir/level{L}_{params,baseMat,matAssignments,elasticityTensors}.bin.
// Returns the number of assignments; throws HostError on a missing input.
int runTensorHomogenization(const std::string &dir, int level, int nbBlockRep, int nbSubdiv,
                            const std::vector<BaseMaterial> &baseMaterials,
                            const std::vector<double> &baseMaterialDensities, bool cubicOnly,
                            const femh_opts *solverOptions = nullptr, int batch = 16, int streams = 3);

// ---- thermal expansion ---------------------------------------------------------
// ThermalAnalysis (SamplesGenerator/ThermalAnalysis.h; run3D :210-274,
// computeThermalCoefficient :388-505) on the engine: per assignment, C_H and
// the effective thermal expansion of the periodic cell, with material 1
// expanding by 1 per unit temperature rise and the other materials not at all
// (the reference's heat forces come from material 1 only, :404-420).  The
// other materials take m_densityMin * E_solid with the solid's Poisson ratio,
// as the multigrid coarsening does.  The expansion comes from the six elastic
// fields with no thermal solve, so iParameters and iTensors (inputs of the
// reference's coarse-element estimate) are not needed.
class ThermalAnalysis {
 public:
  ThermalAnalysis();
  void setSolverOptions(const femh_opts &o) { m_opts = o; }
  // appends, per assignment, the mean normal expansion (a_xx + a_yy + a_zz) / 3
  // (for a cubic cell, its expansion coefficient); `batch` cells per launch
  // sequence
  bool run3D(int N[3], const std::vector<BaseMaterial> &iBaseMaterials,
             const std::vector<std::vector<int>> &iMaterialAssignments, const std::vector<float> &iParameters,
             const std::vector<float> &iTensors, int iNbBlockRep, int iNbSubdiv,
             std::vector<float> &oThermalCoefficients, int batch = 16);
  // the full expansions (Voigt, engineering shear) of the last run3D, 6 per
  // assignment
  const std::vector<double> &lastExpansions() const { return m_alpha; }

 private:
  femh_opts m_opts;
  std::vector<double> m_alpha;
  double m_densityMin = 1e-4;
};

// SamplesGeneratorImpl::runThermalAnalysis (SamplesGeneratorImpl.cpp:4654-4680)
// as a GPU batch job: reads dir/level{L}_matAssignments.bin and writes
// dir/level{L}_thermalExpansionCoefficients.bin (one float per assignment, as
// ThermalAnalysis::run3D).  Returns the number of assignments.
int runThermalAnalysis(const std::string &dir, int level, int nbBlockRep, int nbSubdiv,
                       const std::vector<BaseMaterial> &baseMaterials, const femh_opts *solverOptions = nullptr,
                       int batch = 16);

// ---- optimiser ---------------------------------------------------------------
class MaterialOptimizer {
 public:
  enum StructureType { General, Cubic, Orthotropic };
  MaterialOptimizer() = default;
  void setStructureType(StructureType t) { m_structureType = t; }
  void setSolverOptions(const femh_opts &o) { m_opts = o; m_haveOpts = true; }

  // Projected gradient descent towards the cubic tensor of (E, nu, G) =
  // iTargetParams[1..3] from the 0/1 assignment iMaterialAssignments (N^3);
  // appends the thresholded assignment after every accepted step.
  bool run3D(int N[3], const std::vector<BaseMaterial> &iBaseMaterials,
             const std::vector<int> &iMaterialAssignments, const std::vector<float> &iCurrentParams,
             const std::vector<float> &iTargetParams, std::vector<std::vector<int>> &oNewMaterialAssignments,
             int nSteps = 30);

  int lineSearch(RealFun *fun, VectorXd &x0, const VectorXd &dir, double &h);
  void gradientDescent(FEM3DFun *fun, VectorXd &x0, int nSteps, std::vector<VectorXd> &oParameters);
  double check_df(RealFun *fun, const VectorXd &x0, double h) { return femh::check_df(fun, x0, h); }
  double infNorm(const VectorXd &a) { return femh::infNorm(a); }
  // isotropic/cubic stiffness in Voigt order with shear modulus G
  static Mat6 computeTargetStrains(int idim, float iYoungModulus, float iPoissonRatio, float iShearModulus);

 private:
  StructureType m_structureType = General;
  femh_opts m_opts{};
  bool m_haveOpts = false;
};

}  // namespace femh
