"""GPU tests of the thermal-expansion homogenization (femh_thermal_expansion):
the effective thermal stress from the six elastic fields against the oracle's
Gauss-point restatement on the same fields, and the effective expansion
against two exact relations: a uniform expansion coefficient expands any cell
uniformly, and Levin's relation for two isotropic phases."""
import numpy as np
import pytest

from fem_amd import FemhError, Homogenizer, cfgbin
from fem_amd import workloads as W
from fem_amd.femh import NO_FIELDS
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NU = 0.3
I6 = np.array([1, 1, 1, 0, 0, 0.0])


def levin(C, E1, E2, phi, nu):
    """Two isotropic phases of equal Poisson ratio, expansion 1 in phase 1 and
    0 in phase 2: alpha_H = phi I + (S_H I - <1/3K> I) / (1/3K1 - 1/3K2)."""
    k = lambda e: e / (3 * (1 - 2 * nu))
    mean_c = phi / (3 * k(E1)) + (1 - phi) / (3 * k(E2))
    return phi * I6 + (np.linalg.solve(C, I6) - mean_c * I6) / (1 / (3 * k(E1)) - 1 / (3 * k(E2)))


def test_thermal_stress_matches_oracle_on_engine_fields():
    n = 16
    E = W.simp_modulus(W.simp_density(n, 21, 3.0))
    alpha = W.uniform(5, n ** 3).reshape(n, n, n)  # a different coefficient in every voxel
    h = Homogenizer((n, n, n), nu=NU, tol=1e-12)
    C, _ = h.homogenize_modulus(E)
    beta, aeff = h.thermal_expansion(alpha)
    ref = O.thermal_stress(E, NU, alpha, h.fields())
    assert np.abs(beta - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(C @ aeff - beta).max() <= 1e-12 * np.abs(beta).max()
    h.close()


def test_uniform_expansion_and_levin_relation():
    n = 32
    h = Homogenizer((n, n, n), nu=NU, tol=1e-12)
    E = W.simp_modulus(W.simp_density(n, 3, 4.0))
    h.homogenize_modulus(E)
    _, aeff = h.thermal_expansion(np.full(E.shape, 0.7))
    assert np.abs(aeff - 0.7 * I6).max() <= 1e-9
    s = W.binary_cube(n, 9) > 0.5
    E = np.where(s, 1.0, 0.05)
    C, _ = h.homogenize_modulus(E)
    _, aeff = h.thermal_expansion(s.astype(float))
    ref = levin(C, 1.0, 0.05, s.mean(), NU)
    assert np.abs(aeff - ref).max() <= 1e-8, np.abs(aeff - ref).max()
    h.close()


def test_thermal_batch_and_cubic_structure():
    """A batch gives the one-cell values; in the cubic-structure mode the four
    fields formed by axis swaps give the same thermal stress as solved ones
    (a quantity first order in the fields, so it checks them directly)."""
    n = 16
    cells = [W.cubic_average(W.simp_modulus(W.simp_density(n, s, 3.0))) for s in (1, 2)]
    cells.append(W.simp_modulus(W.simp_density(n, 7, 3.0)))  # not cubic
    Es = np.stack(cells)
    alphas = np.stack([W.cubic_average(W.uniform(s, n ** 3).reshape(n, n, n)) for s in (11, 12, 13)])
    hb = Homogenizer((n, n, n), nu=NU, tol=1e-12, batch=3, structure="cubic")
    hb.homogenize_many(Es, modulus=True)
    bb, ab = hb.thermal_expansion(alphas, nb=3)
    h1 = Homogenizer((n, n, n), nu=NU, tol=1e-12)
    for s in range(3):
        h1.homogenize_modulus(Es[s])
        b1, a1 = h1.thermal_expansion(alphas[s])
        assert np.abs(bb[s] - b1).max() <= 1e-9 * np.abs(b1).max(), s
        assert np.abs(ab[s] - a1).max() <= 1e-9 * np.abs(a1).max(), s
    hb.close()
    h1.close()


def test_thermal_needs_a_solve():
    h = Homogenizer((8, 8, 8), nu=NU)
    with pytest.raises(FemhError) as e:
        h.thermal_expansion(np.ones((8, 8, 8)))
    assert e.value.status == NO_FIELDS
    h.close()


def test_run_thermal_analysis_file_job(tmp_path):
    level, nsamp = 4, 5
    rng = np.random.default_rng(23)
    assign = [rng.integers(0, 2, level ** 3).astype(np.int32) for _ in range(nsamp)]
    assign[0][:] = 1
    root = tmp_path / ("level%d_" % level)
    cfgbin.write_vectors(str(root) + "matAssignments.bin", assign, np.int32)
    coef, alpha = cfgbin.run_thermal_analysis(str(tmp_path), level, 2.0, 0.35, subdiv=2, batch=2, tol=1e-12)
    back = cfgbin.read_vector(str(root) + "thermalExpansionCoefficients.bin", np.float32)
    np.testing.assert_array_equal(back, coef)
    assert abs(coef[0] - 1.0) < 1e-6  # all solid
    n = 2 * level
    h = Homogenizer((n, n, n), nu=0.35, E0=2.0, interp="modulus", tol=1e-12)
    for s in range(1, nsamp):
        solid = cfgbin.replicate_assignment(assign[s], level, 1, 2) == 1
        C, _ = h.homogenize_modulus(2.0 * np.where(solid, 1.0, 1e-4))
        _, a1 = h.thermal_expansion(solid.astype(float))
        assert np.abs(alpha[s] - a1).max() <= 1e-9, s
        lev = levin(C, 2.0, 2e-4, solid.mean(), 0.35)
        assert np.abs(a1 - lev).max() <= 1e-6, (s, np.abs(a1 - lev).max())
    h.close()

