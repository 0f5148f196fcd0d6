"""GPU tests of the cubic-structure fast path (femh_opts.structure): on cells
invariant under the 48 axis permutations and reflections, solving the xx and
yz load cases gives the full C_H of the six-case solve."""
import numpy as np
import pytest

from fem_amd import FemhError, Homogenizer
from fem_amd import workloads as W
from fem_amd.femh import BAD_ARG

pytestmark = pytest.mark.gpu
NU = 0.3


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cells(n, seeds):
    out = []
    for s in seeds:
        rho = W.cubic_average(W.simp_density(n, s, 3.0))
        out.append(rho ** 3)  # SIMP moduli
        b = W.cubic_average(W.binary_cube(n, s + 100).astype(float)) > 0.5
        out.append(np.where(b, 1.0, 1e-4))  # binary with a soft phase
    return out


def test_cubic_fp64_equals_six_case_solve():
    """C_H and the adjoint from two solves plus axis swaps equal those of the
    six-case solve."""
    hc = Homogenizer((16, 16, 16), nu=NU, tol=1e-12, structure="cubic")
    hg = Homogenizer((16, 16, 16), nu=NU, tol=1e-12)
    G = np.random.default_rng(4).standard_normal((6, 6))
    for E in cells(16, [1, 2]):
        Cc, st = hc.homogenize_modulus(E)
        Cg, _ = hg.homogenize_modulus(E)
        assert rel(Cc, Cg) <= 1e-9, rel(Cc, Cg)
        assert [st["iters"][k] for k in (1, 2, 4, 5)] == [0, 0, 0, 0]
        assert st["iters"][0] > 0 and st["iters"][3] > 0
        gc, gg = hc.gradient(G), hg.gradient(G)
        assert np.abs(gc - gg).max() <= 1e-8 * np.abs(gg).max()


def test_cubic_mode_solves_asymmetric_cells_in_full():
    """In a cubic-mode batch, a cell that is not invariant under the axis swaps
    takes all six solves; its neighbours keep the two-solve path."""
    E = cells(16, [3])
    asym = W.simp_modulus(W.simp_density(16, 8, 3.0))
    Es = np.stack([E[0], asym, E[1]])
    hc = Homogenizer((16, 16, 16), nu=NU, tol=1e-12, structure="cubic", batch=3)
    hg = Homogenizer((16, 16, 16), nu=NU, tol=1e-12, batch=3)
    Cc, stc = hc.homogenize_many(Es, modulus=True)
    Cg, _ = hg.homogenize_many(Es, modulus=True)
    for s in range(3):
        assert rel(Cc[s], Cg[s]) <= 1e-9, (s, rel(Cc[s], Cg[s]))
    assert [stc[s]["iters"][1] > 0 for s in range(3)] == [False, True, False]
    assert abs(Cg[1][0, 3]) > 1e-6 * Cg[1][0, 0]  # an anisotropic coupling the average would erase


def test_cubic_needs_a_cube_grid():
    with pytest.raises(FemhError) as e:
        Homogenizer((16, 16, 8), nu=NU, structure="cubic")
    assert e.value.status == BAD_ARG


def test_cubic_mixed_batch_within_1e6():
    Es = np.stack(cells(32, [5, 6]))
    hm = Homogenizer((32, 32, 32), nu=NU, precision="mixed", ch_tol=1e-7, inner_tol=3e-4, batch=4,
                     structure="cubic")
    Cm, stm = hm.homogenize_many(Es, modulus=True)
    hg = Homogenizer((32, 32, 32), nu=NU, tol=1e-12, batch=4)
    Cg, _ = hg.homogenize_many(Es, modulus=True)
    for s in range(len(Es)):
        assert rel(Cm[s], Cg[s]) <= 1e-6, (s, rel(Cm[s], Cg[s]))
        assert stm[s]["iters"][1] == 0 and stm[s]["iters"][5] == 0
