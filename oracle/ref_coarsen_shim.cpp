// ref_coarsen_shim.cpp -- C-ABI entry into the reference's own femlib code of
// the force-driven homogenization path, compiled from the sources under the
// reference tree (see oracle/Makefile, target `ref`).  TEST INFRASTRUCTURE
// ONLY: tests/ use it to pin the oracle's restatement of that path; nothing
// here solves a system.
//
// ref_constraint_triplets: the constraint rows femlib appends to a stiffness
//   matrix -- ElementMesh::fixTranslation, fixRotation and enforcePeriodicity
//   (femlib/src/EMeshStiffness.cpp:53-274) in the lower-triangular form
//   NumericalCoarsening::getStiffnessSparse asks for
//   (SamplesGenerator/NumericalCoarsening.cpp:1283-1291) -- as (row, col,
//   value) triplets with row >= 3 * vertices.
// ref_coarse_tensor: meshUtil::computeCoarsenedElasticityTensor
//   (femlib/src/MeshUtilities.cpp:960-1060) of six given vertex displacement
//   fields, on an ElementRegGrid whose element e carries a MaterialQuad over
//   StrainLin with the Lame parameters of modulus E[e] and Poisson ratio nu.
#include <memory>
#include <vector>

#include "ElementRegGrid.hpp"
#include "MaterialQuad.hpp"
#include "MeshUtilities.h"
#include "StrainLin.hpp"

extern "C" int ref_constraint_triplets(int nx, int ny, int nz, int max, int *rows, int *cols, double *vals) {
  ElementRegGrid grid(nx, ny, nz);
  const int N = 3 * (int)grid.x.size();
  Eigen::SparseMatrix<cfgScalar> K(N, N);
  grid.fixTranslation(K, true, &grid);
  grid.fixRotation(K, true, &grid);
  grid.enforcePeriodicity(K, true, &grid);
  int n = 0;
  for (int c = 0; c < K.outerSize(); c++)
    for (Eigen::SparseMatrix<cfgScalar>::InnerIterator it(K, c); it; ++it) {
      if (it.value() == 0) continue;
      if (n < max) {
        rows[n] = (int)it.row();
        cols[n] = (int)it.col();
        vals[n] = (double)it.value();
      }
      n++;
    }
  return n;
}

extern "C" int ref_coarse_tensor(int nx, int ny, int nz, const double *E, double nu, const double *u, double *C) {
  ElementRegGrid grid(nx, ny, nz);
  const int nel = nx * ny * nz, nd = 3 * (int)grid.X.size();
  std::vector<std::unique_ptr<StrainLin>> ene;
  std::vector<std::unique_ptr<MaterialQuad>> mats;
  for (int e = 0; e < nel; e++) {
    ene.emplace_back(new StrainLin());
    const double lam = E[e] * nu / ((1 + nu) * (1 - 2 * nu)), mu = E[e] / (2 * (1 + nu));
    ene.back()->param[0] = (cfgScalar)mu;
    ene.back()->param[1] = (cfgScalar)lam;
    mats.emplace_back(new MaterialQuad(ene.back().get()));
    grid.addMaterial(mats.back().get());
    grid.me[e] = e;
  }
  std::vector<std::vector<cfgScalar>> disp(6, std::vector<cfgScalar>(nd));
  for (int d = 0; d < 6; d++)
    for (int t = 0; t < nd; t++) disp[d][t] = (cfgScalar)u[(size_t)d * nd + t];
  MatrixXS Ch;
  meshUtil::computeCoarsenedElasticityTensor(grid, disp, Ch);
  if (Ch.rows() != 6 || Ch.cols() != 6) return 1;
  for (int a = 0; a < 6; a++)
    for (int b = 0; b < 6; b++) C[6 * a + b] = (double)Ch(a, b);
  return 0;
}
