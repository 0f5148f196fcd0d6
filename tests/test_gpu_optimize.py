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


@pytest.mark.parametrize("mw", [0.0, 0.5])
def test_device_loop_matches_the_host_loop(mw):
    """DeviceDescent (densities, gradients and trial points in device memory,
    femh_gradient_device / femh_descent_trial) against BatchedDescent (the
    same loop on host arrays) over 3 steps of 5 cells in groups of 2 (the last
    group partial), mixed precision: without a volume term the trial points
    are the same numbers, so the trajectories are identical; with one the
    cells' means differ only in summation order."""
    from fem_amd.optimize import DeviceDescent
    n, cells = 12, 5
    x0 = np.stack([W.simp_density(n, 20 + s, 2.5).ravel() for s in range(cells)])
    C0 = np.stack([iso(0.12 + 0.04 * s, 0.28) for s in range(cells)])
    kw = dict(m0=0.35, mw=mw, batch=2, Emin=1e-3, precision="mixed", tol=1e-10, ch_tol=1e-7)
    hd = BatchedDescent((n, n, n), C0, **kw)
    dd = DeviceDescent((n, n, n), C0, **kw)
    xh, xd = x0.copy(), dd.to_device(x0)
    for k in range(3):
        xh, Jh, ah, J0h = hd.step(xh, 0.6)
        xd, Jd, ad, J0d = dd.step(xd, 0.6)
        xdh = xd.cpu().numpy()
        assert np.array_equal(ah, ad), k
        assert ah.any(), "the steps move"
        if mw == 0.0:
            assert np.array_equal(xh, xdh), k
            assert np.array_equal(Jh, Jd) and np.array_equal(J0h, J0d), k
        else:
            assert np.abs(xh - xdh).max() <= 1e-12, k
            assert np.abs(Jh - Jd).max() <= 1e-12 * np.abs(Jh).max(), k
        assert np.abs(dd.gmax - np.abs(dd.G.cpu().numpy()).max(axis=1)).max() == 0.0
    assert hd.solves == dd.solves
    hd.close()
    dd.close()


def test_device_trial_rounding_and_modes():
    """femh_descent_trial per cell: h > 0 gives clamp(x - h g) with the product
    and the difference rounded separately (bit-equal to numpy), h < 0 copies
    x, h == 0 leaves the output; the sums are the cells' sums."""
    import torch
    n = 8
    h = Homogenizer((n, n, n), nu=0.3)
    rng = np.random.default_rng(3)
    V = n ** 3
    x = rng.uniform(0.0, 1.0, (3, V))
    g = rng.standard_normal((3, V)) * 3.0
    out0 = rng.uniform(0.0, 1.0, (3, V))
    dx, dg, do = (torch.from_numpy(a.copy()).cuda() for a in (x, g, out0))
    torch.cuda.synchronize()
    hs = np.array([0.137, -1.0, 0.0])
    sums = h.descent_trial(3, dx.data_ptr(), dg.data_ptr(), hs, 1e-3, 1.0, do.data_ptr())
    res = do.cpu().numpy()
    assert np.array_equal(res[0], np.clip(x[0] - hs[0] * g[0], 1e-3, 1.0))
    assert np.array_equal(res[1], x[1])
    assert np.array_equal(res[2], out0[2])
    assert np.allclose(sums, res.sum(axis=1), rtol=1e-14, atol=0)
    h.close()
