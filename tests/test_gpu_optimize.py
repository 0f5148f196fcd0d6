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
