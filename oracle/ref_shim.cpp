// ref_shim.cpp -- C-ABI entry into the reference's own femlib element code,
// compiled from the sources under the reference tree (see oracle/Makefile).
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's Ke against the
// reference implementation itself.
//
// ref_ke_hex builds a one-voxel ElementRegGrid (femlib/src/ElementRegGrid.cpp:
// 107-142, dx = 1/max(n) = 1), a MaterialQuad over StrainLin with
// param = {mu, lambda} (femlib/src/StrainLin.cpp:4-9), and returns
// MaterialQuad::getStiffness (femlib/src/MaterialQuad.cpp:68-79) row-major.
#include "ElementRegGrid.hpp"
#include "MaterialQuad.hpp"
#include "StrainLin.hpp"
#include "Element.hpp"

extern "C" int ref_ke_hex(double mu, double lambda, double* out) {
  ElementRegGrid grid(1, 1, 1);
  StrainLin ene;
  ene.param[0] = (cfgScalar)mu;
  ene.param[1] = (cfgScalar)lambda;
  MaterialQuad mat(&ene);
  grid.addMaterial(&mat);
  mat.init(&grid);
  MatrixXS K = mat.getStiffness(grid.e[0], &grid);
  if (K.rows() != 24 || K.cols() != 24) return 1;
  for (int r = 0; r < 24; r++)
    for (int c = 0; c < 24; c++) out[24 * r + c] = (double)K(r, c);
  return 0;
}

extern "C" int ref_scalar_bytes(void) { return (int)sizeof(cfgScalar); }
