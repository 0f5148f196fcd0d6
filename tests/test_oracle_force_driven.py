"""CPU pins of the oracle's force-driven path (the reference's NumericalCoarsening
solve: face tractions, periodic / translation / rotation constraints on the
(n+1)^3-vertex mesh, coarse-strain tensor of femlib/src/MeshUtilities.cpp:
960-1060), and of the identity the engine's coarse-strain tensor rests on.

* Against the reference's own femlib code, compiled from its sources into
  oracle/_ref/libref_coarsen_double.so (skipped where that library is not
  built): the constraint rows equal those of ElementMesh::fixTranslation /
  fixRotation / enforcePeriodicity as a set, and meshUtil::
  computeCoarsenedElasticityTensor of the oracle's displacement fields equals
  the oracle's tensor.  PARDISO, the reference's solver, is closed and absent:
  the oracle solves the same system by dense LU.
* By the mathematics: a homogeneous cell gives the isotropic tensor; on cube
  cells, without the 1e-5 diagonal regularisation, the constrained solutions
  are macro strains plus strain-driven fluctuations, so the tensor equals the
  formula on the strain-driven fields (or_coarse_strain_from_fields); the
  regularisation moves it by far less than 1e-5.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REFLIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_coarsen_double.so")
NU = 0.3


def c_iso(E, nu):
    lam, mu = O.lame(E, nu)
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _ref():
    if not os.path.exists(REFLIB):
        pytest.skip("reference femlib coarsening library not built (needs the reference tree)")
    return ctypes.CDLL(REFLIB)


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 3, 3), (3, 4, 5)])
def test_constraint_rows_equal_femlib(shape):
    L = _ref()
    nx, ny, nz = shape
    ND = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    cap = 400000
    r = np.zeros(cap, np.int32)
    c = np.zeros(cap, np.int32)
    v = np.zeros(cap)
    vp = ctypes.c_void_p
    n = L.ref_constraint_triplets(nx, ny, nz, cap, vp(r.ctypes.data), vp(c.ctypes.data), vp(v.ctypes.data))
    assert 0 < n <= cap
    r, c, v = r[:n], c[:n], v[:n]
    assert (r >= ND).all()  # lower-triangular form: constraint rows only
    ref = np.zeros((r.max() - ND + 1, ND))
    np.add.at(ref, (r - ND, c), v)
    ours = O.kkt_constraints(shape, faithful=True)
    assert ref.shape == ours.shape
    assert sorted(map(tuple, ref)) == sorted(map(tuple, ours))


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 3, 3), (3, 4, 5)])
def test_coarse_tensor_equals_femlib_reduction(shape):
    L = _ref()
    E = np.random.default_rng(2).uniform(0.05, 1.0, shape)
    fd = O.force_driven(E, NU, regularize=True, faithful=True)
    assert fd["status"] == 0
    C = np.zeros(36)
    vp = ctypes.c_void_p
    st = L.ref_coarse_tensor(*shape, vp(E.ctypes.data), ctypes.c_double(NU), vp(fd["u"].ctypes.data),
                             vp(C.ctypes.data))
    assert st == 0
    assert relmax(C.reshape(6, 6), fd["C"]) <= 1e-13


@pytest.mark.parametrize("regularize", [False, True])
def test_homogeneous_cell_gives_isotropic_tensor(regularize):
    fd = O.force_driven(np.full((3, 3, 3), 2.0), NU, regularize=regularize)
    assert relmax(fd["C"], c_iso(2.0, NU)) <= (1e-13 if not regularize else 1e-5)


@pytest.mark.parametrize("shape,seed", [((2, 2, 2), 1), ((3, 3, 3), 2), ((4, 4, 4), 3)])
def test_cube_cells_equal_the_strain_driven_formula(shape, seed):
    E = np.random.default_rng(seed).uniform(0.02, 1.0, shape)
    fd = O.force_driven(E, NU)
    sd = O.homogenize(E, NU, precond="jacobi", tol=1e-14, max_iter=50000)
    Cf = O.coarse_strain_from_fields(E, NU, sd["u"])
    assert relmax(fd["C"], Cf) <= 1e-12
    # the convention (engineering stiffness on tensor shear strains) differs
    # from the energy form on a heterogeneous cell
    if shape != (2, 2, 2):
        assert relmax(fd["C"], sd["C"]) > 1e-3
    # the reference's diagonal regularisation moves it by far less than 1e-5
    fr = O.force_driven(E, NU, regularize=True)
    assert relmax(fr["C"], fd["C"]) < 1e-5
    if shape[0] >= 3:
        assert relmax(fr["C"], fd["C"]) > 1e-9


def test_non_cube_cell_shear_loads_carry_a_moment():
    """On a non-cube cell the reference's shear tractions (equal totals on
    faces of different size) carry a net moment that the rotation constraint
    takes up: the result is then not the strain-driven formula's."""
    E = np.random.default_rng(4).uniform(0.05, 1.0, (4, 3, 5))
    fd = O.force_driven(E, NU)
    sd = O.homogenize(E, NU, precond="jacobi", tol=1e-14, max_iter=50000)
    assert relmax(fd["C"], O.coarse_strain_from_fields(E, NU, sd["u"])) > 1e-4
