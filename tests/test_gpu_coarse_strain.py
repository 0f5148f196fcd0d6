"""GPU tests of femh_coarse_strain_tensor, the force-driven (coarse-strain)
form of C_H: against the oracle's formula on the same strain-driven fields,
against the oracle's restatement of the reference's force-driven KKT solve
(pinned to the reference's femlib code in tests/test_oracle_force_driven.py),
exactly isotropic on a homogeneous cell, batched and mixed precision, and
rejected on a non-cube cell."""
import numpy as np
import pytest

from fem_amd import FemhError, Homogenizer
from fem_amd import workloads as W
from fem_amd.femh import BAD_ARG, NO_FIELDS
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NU = 0.3


def relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


@pytest.mark.parametrize("n", [4, 8, 12])
def test_matches_oracle_formula_on_oracle_fields(n):
    E = W.simp_modulus(W.simp_density(n, 9, 2.0))
    h = Homogenizer((n, n, n), nu=NU, tol=1e-12, interp="modulus")
    h.homogenize_modulus(E)
    Cg = h.coarse_strain_tensor()
    ref = O.homogenize(E, NU, precond="mg", tol=1e-13)
    Co = O.coarse_strain_from_fields(E, NU, ref["u"])
    assert relmax(Cg, Co) <= 1e-10
    assert np.abs(Cg - Cg.T).max() <= 1e-14 * np.abs(Cg).max()


@pytest.mark.parametrize("n", [3, 4])
def test_matches_reference_force_driven_solve(n):
    """The engine's tensor from its strain-driven solve equals the oracle's
    dense force-driven KKT solve of the same cell (no diagonal
    regularisation)."""
    E = np.random.default_rng(5 + n).uniform(0.02, 1.0, (n, n, n))
    h = Homogenizer((n, n, n), nu=NU, tol=1e-13, precond="jacobi", max_iter=20000, interp="modulus")
    h.homogenize_modulus(E)
    fd = O.force_driven(E, NU)
    assert fd["status"] == 0
    assert relmax(h.coarse_strain_tensor(), fd["C"]) <= 1e-9


def test_homogeneous_cell_and_errors():
    h = Homogenizer((6, 6, 6), nu=NU, interp="modulus")
    with pytest.raises(FemhError) as e:
        h.coarse_strain_tensor()
    assert e.value.status == NO_FIELDS
    h.homogenize_modulus(np.full((6, 6, 6), 3.0))
    lam, mu = O.lame(3.0, NU)
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    assert np.abs(h.coarse_strain_tensor() - C).max() <= 1e-13
    hr = Homogenizer((6, 6, 8), nu=NU, interp="modulus")
    hr.homogenize_modulus(np.full((6, 6, 8), 1.0))
    with pytest.raises(FemhError) as e:
        hr.coarse_strain_tensor()
    assert e.value.status == BAD_ARG


def test_batched_and_mixed_precision():
    n = 32
    rho = np.stack([W.simp_density(n, 50 + s, 4.0) for s in range(3)])
    h64 = Homogenizer((n, n, n), nu=NU, tol=1e-11, batch=3)
    h64.homogenize_many(rho)
    C64 = h64.coarse_strain_tensor(nb=3)
    h1 = Homogenizer((n, n, n), nu=NU, tol=1e-11)
    for s in range(3):
        h1.homogenize(rho[s])
        assert relmax(C64[s], h1.coarse_strain_tensor()) <= 1e-12
    # its error is first order in the field error (the form is not the
    # solve's energy), so mixed precision refines to the residual tolerance
    hm = Homogenizer((n, n, n), nu=NU, precision="mixed", tol=1e-11, ch_tol=0.0, batch=3)
    hm.homogenize_many(rho)
    Cm = hm.coarse_strain_tensor(nb=3)
    for s in range(3):
        assert np.linalg.norm(Cm[s] - C64[s]) / np.linalg.norm(C64[s]) <= 1e-8
