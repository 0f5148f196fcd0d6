"""fem_amd -- MI355X (gfx950) voxel-FEM periodic homogenization engine.

Python face of the C ABI in include/femh.h (built as fem_amd/libfemh.so):
density grid in, homogenized elasticity tensor C_H (6x6, Voigt xx,yy,zz,yz,xz,xy,
engineering shear) and dJ/drho out.
"""
from .femh import (BatchHomogenizer, FemhError, Homogenizer, homogenize_batch, lib,  # noqa: F401
                   LIB_PATH, EXPORTS)

__all__ = ["BatchHomogenizer", "FemhError", "Homogenizer", "homogenize_batch", "lib", "LIB_PATH",
           "EXPORTS"]
