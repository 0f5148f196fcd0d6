"""The batched optimisation workload at its full size: 64 SIMP cells of 96^3
(E = 1e-3 + (1 - 1e-3) rho^3, MaterialOpt/src/FEM3DFun.cpp:244-257), each
matched to its own seeded isotropic target C_H, by projected-gradient steps
with a halving line search (fast_fem/run3d.cpp:238-289) on warm-started
mixed-precision multigrid-PCG, 16 cells per launch sequence, with the
densities, gradients and trial points in device memory (DeviceDescent).  Two
full steps are run; at the start point of each step, and at the final point, the C_H of
every cell is checked against the FP64 path (<= 1e-6 relative), and every
cell's objective must not increase."""
import numpy as np
import pytest

from fem_amd import Homogenizer
from fem_amd import workloads as W
from fem_amd.optimize import DeviceDescent
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def iso(E, nu):
    lam, mu = O.lame(E, nu)
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_opt96_two_steps_per_step_error_vs_fp64():
    n, count = 96, 64
    x = np.stack([W.simp_density(n, 100 + s, 6.0).reshape(-1) for s in range(count)])
    Et = 0.05 + 0.1 * W.uniform(5, count)
    C0 = np.stack([iso(float(e), 0.25) for e in Et])
    bd = DeviceDescent((n, n, n), C0, batch=16, precision="mixed", tol=1e-10, ch_tol=1e-7,
                       inner_tol=3e-4, Emin=1e-3)
    ref = Homogenizer((n, n, n), nu=0.3, tol=1e-10, batch=8, Emin=1e-3)
    x = bd.to_device(x)
    try:
        for k in range(3):
            xh = x.cpu().numpy()
            if k < 2:
                xn, Jn, acc, J0 = bd.step(x)
                Cm = bd.last_C
            else:
                Cm, _ = bd.evaluate(x)
            C64, _ = ref.homogenize_many(xh.reshape(count, n, n, n))
            errs = np.array([rel(Cm[s], C64[s]) for s in range(count)])
            print(f"step {k}: C_H error vs FP64 over {count} cells: max {errs.max():.2e}, "
                  f"median {np.median(errs):.2e}")
            assert errs.max() <= 1e-6, (k, int(errs.argmax()), errs.max())
            if k < 2:
                assert (Jn <= J0).all()
                assert acc.sum() >= count // 2
                x = xn
    finally:
        bd.close()
        ref.close()
