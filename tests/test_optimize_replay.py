"""Replay pin of the batched optimiser (fem_amd/optimize.py) on the CPU oracle.

The per-cell loop is fast_fem/run3d.cpp:238-289: gradientDescent takes the
gradient at the current point, h = maxStep / ||g||_inf, and lineSearch tries
x - h g clamped to the bounds, halving h up to 10 times until the objective
drops (and stops the descent when it never does).  The objective and its
gradient are those of fast_fem/FEM3DFun.cpp:198-268 with the SIMP modulus of
MaterialOpt/src/FEM3DFun.cpp:244-257.

Here that loop is restated literally for one cell (`reference_descent`,
driven by the oracle's C_H and adjoint) and compared with the batched loop
(`BatchedDescent.step`, which advances all cells at once with per-cell
line-search bookkeeping) driven by the same oracle: the trajectories must be
identical.  A GPU test (tests/test_gpu_optimize.py) then compares the engine-
driven batched loop with this oracle-driven one.
"""
import numpy as np
import pytest

from fem_amd import workloads as W
from fem_amd.optimize import BatchedDescent, objective
from oracle import oracle as O

NU = 0.3
EMIN = 1e-3
PENAL = 3.0


def iso(E, nu):
    lam, mu = O.lame(E, nu)
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def modulus(x):
    return EMIN + (1.0 - EMIN) * x ** PENAL


class OracleCell:
    """One cell as the reference's RealFun (setParam / f / df) on the oracle."""

    def __init__(self, shape, C0, m0, mw):
        self.shape, self.C0, self.m0, self.mw = shape, C0, m0, mw

    def setParam(self, x):
        self.x = x.copy()
        r = O.homogenize(modulus(x).reshape(self.shape), NU, precond="mg", tol=1e-13)
        self.C, self.u = r["C"], r["u"]

    def f(self):
        return float(objective(self.C, self.C0, np.ones((6, 6)), self.x.mean(), self.m0, self.mw))

    def df(self):
        dE = PENAL * (1.0 - EMIN) * self.x ** (PENAL - 1.0)
        g = dE * O.grad_E(self.shape, NU, self.u, self.C - self.C0).ravel()
        return g + self.mw * (self.x.mean() - self.m0) / self.x.size


def reference_descent(fun, x0, nsteps, max_step, lb=1e-3, ub=1.0):
    """run3d.cpp:238-289 for one cell: returns the point after every step."""
    x = x0.copy()
    fun.setParam(x)
    traj = [x.copy()]
    for _ in range(nsteps):
        grad = fun.df()
        h = max_step / np.abs(grad).max()
        f0 = fun.f()
        ret = -1
        for _ in range(10):
            xt = np.clip(x - h * grad, lb, ub)
            fun.setParam(xt)
            if fun.f() < f0:
                x = xt
                ret = 0
                break
            h /= 2
        if ret < 0:
            break
        traj.append(x.copy())
    return traj


class OracleDescent(BatchedDescent):
    """The batched loop with the oracle as its solver (one group, CPU)."""

    def _open(self, nu, interp, penal, solver):
        assert interp == "simp" and penal == PENAL and solver.get("Emin") == EMIN
        self._fields = {}
        return []

    def _group_solve(self, gi, xg):
        C = np.zeros((len(xg), 6, 6))
        for k, c in enumerate(self.groups[gi]):
            r = O.homogenize(modulus(xg[k]).reshape(self.shape), NU, precond="mg", tol=1e-13)
            C[k] = r["C"]
            self._fields[c] = (xg[k].copy(), r["u"])
        return C, [{"ch_err_est": 0.0} for _ in xg]

    def _group_gradient(self, gi, dJdC):
        g = np.zeros((len(dJdC), int(np.prod(self.shape))))
        for k, c in enumerate(self.groups[gi]):
            xk, u = self._fields[c]
            dE = PENAL * (1.0 - EMIN) * xk ** (PENAL - 1.0)
            g[k] = dE * O.grad_E(self.shape, NU, u, dJdC[k]).ravel()
        return g

    def close(self):
        self._pool.shutdown()


def problem(n=6, cells=3):
    x0 = np.stack([W.simp_density(n, 40 + s, 2.0).ravel() for s in range(cells)])
    C0 = np.stack([iso(0.15 + 0.05 * s, 0.25) for s in range(cells)])
    return (n, n, n), x0, C0


def test_batched_loop_replays_the_reference_loop_on_the_oracle():
    shape, x0, C0 = problem()
    m0, mw, steps, max_step = 0.4, 0.5, 4, 1.0  # 2-3 halvings per step and cell
    od = OracleDescent(shape, C0, m0=m0, mw=mw, batch=len(x0), Emin=EMIN)
    x = x0.copy()
    traj = [x.copy()]
    for _ in range(steps):
        x, J, acc, J0 = od.step(x, max_step)
        assert acc.all(), "every line search succeeds on this problem"
        assert (J < J0).all()
        traj.append(x.copy())
    od.close()
    for s in range(len(x0)):
        ref = reference_descent(OracleCell(shape, C0[s], m0, mw), x0[s], steps, max_step)
        assert len(ref) == steps + 1
        for k in range(steps + 1):
            np.testing.assert_allclose(traj[k][s], ref[k], rtol=0, atol=1e-14)


def test_reference_loop_stops_when_the_line_search_fails():
    """A full cell (every density at the upper bound) asked to be twice as
    stiff: the gradient points outside the bounds, every clamped trial is the
    start point and no step lowers the objective.  The reference stops there;
    the batched step reports the cell as not accepted and leaves it in place."""
    shape = (4, 4, 4)
    x0 = np.ones((1, 64))
    C0 = 2.0 * iso(1.0, NU)[None]
    fun = OracleCell(shape, C0[0], 0.0, 0.0)
    fun.setParam(x0[0])
    assert (fun.df() < 0).all()
    ref = reference_descent(fun, x0[0], 3, 0.3)
    assert len(ref) == 1
    od = OracleDescent(shape, C0, batch=1, Emin=EMIN)
    x1, J1, acc, J0 = od.step(x0, 0.3)
    od.close()
    assert not acc[0] and np.array_equal(x1, x0) and J1[0] == J0[0]
