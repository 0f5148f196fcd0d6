"""GPU test of the batched projected-gradient loop (fem_amd/optimize.py):
groups of cells solved at once give the same C_H and gradients as separate
solves, and steps lower every cell's objective."""
import numpy as np
import pytest

from fem_amd import Homogenizer
from fem_amd import workloads as W
from fem_amd.optimize import BatchedDescent

pytestmark = pytest.mark.gpu


def iso(E, nu):
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[range(3, 6), range(3, 6)] = mu
    return C


def test_batched_descent_groups_and_steps():
    n, cells = 16, 4
    x = np.stack([W.simp_density(n, s, 3.0).ravel() for s in range(cells)])
    C0 = np.stack([iso(0.2 + 0.05 * s, 0.3) for s in range(cells)])
    bd = BatchedDescent((n, n, n), C0, batch=2, Emin=1e-3, tol=1e-10)
    C, G = bd.evaluate(x, grad=True)
    # the same cells through one context each, one at a time
    for s in range(cells):
        h = Homogenizer((n, n, n), nu=0.3, Emin=1e-3, tol=1e-10)
        Cs, _ = h.homogenize(x[s].reshape(n, n, n))
        gs = h.gradient(bd.wG[s] * (Cs - C0[s])).ravel()
        h.close()
        assert np.abs(C[s] - Cs).max() <= 1e-9 * np.abs(Cs).max()
        assert np.abs(G[s] - gs).max() <= 1e-6 * np.abs(gs).max()
    J0 = bd.J(C, x)
    x1, J1, acc, _ = bd.step(x, max_step=0.2)
    assert acc.all() and (J1 < J0).all()
    bd.close()


def test_engine_driven_loop_matches_the_oracle_replay():
    """The batched loop driven by the engine (FP64, warm-started batch of 3)
    follows the oracle-driven one, which tests/test_optimize_replay.py pins to
    a literal restatement of fast_fem/run3d.cpp:238-289: same accepted steps,
    same points to solver tolerance."""
    from test_optimize_replay import EMIN, OracleDescent, problem
    shape, x0, C0 = problem()
    kw = dict(m0=0.4, mw=0.5, batch=len(x0), Emin=EMIN)
    gd = BatchedDescent(shape, C0, tol=1e-12, **kw)
    od = OracleDescent(shape, C0, **kw)
    xg, xo = x0.copy(), x0.copy()
    for _ in range(4):
        xg, Jg, ag, _ = gd.step(xg, 1.0)
        xo, Jo, ao, _ = od.step(xo, 1.0)
        assert np.array_equal(ag, ao)
        assert np.abs(xg - xo).max() <= 1e-9
        assert np.abs(Jg - Jo).max() <= 1e-9 * np.abs(Jo).max()
    gd.close()
    od.close()
