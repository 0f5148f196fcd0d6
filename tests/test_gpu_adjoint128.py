"""Adjoint at 128^3 (GRF correlation 8, SIMP p = 3): dJ/drho for
J = ||C_H - C_t||_F^2, C_t the isotropic tensor of E = 0.3, nu = 0.25, against
central differences (h = 1e-6) on 64 seeded voxels.  The 128 perturbed cells
run 8 per launch sequence, each warm-started from the previous batch's fields."""
import numpy as np
import pytest

from fem_amd import Homogenizer
from fem_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NU = 0.3


def iso(E, nu):
    lam, mu = O.lame(E, nu)
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_128_matches_central_differences(precision):
    """precision "mixed": FP32 multigrid-PCG corrections inside FP64
    iterative refinement to the same residual target, the adjoint formed in
    FP64 from those fields."""
    n, hstep = 128, 1e-6
    rho = W.simp_density(n, 4, 8.0)
    Ct = iso(0.3, 0.25)
    h = Homogenizer((n, n, n), nu=NU, penal=3.0, tol=1e-12, max_iter=4000, batch=8,
                    warm_start=True, precision=precision)
    C0, st = h.homogenize(rho)
    g = h.gradient(2.0 * (C0 - Ct))
    idx = (W.uniform(12, 64) * rho.size).astype(int)
    cells = []
    for f in idx:
        for sgn in (1.0, -1.0):
            r = rho.copy().reshape(-1)
            r[f] += sgn * hstep
            cells.append(r)
    J = []
    for b0 in range(0, len(cells), 8):
        C, _ = h.homogenize_many(np.stack(cells[b0:b0 + 8]))
        J.extend(float(np.sum((c - Ct) ** 2)) for c in C)
    J = np.array(J).reshape(-1, 2)
    fd = (J[:, 0] - J[:, 1]) / (2 * hstep)
    ga = g.reshape(-1)[idx]
    err = np.abs(fd - ga).max() / np.abs(g).max()
    print(f"{precision}: max |fd - adjoint| / max |adjoint| = {err:.3e} over {len(idx)} voxels")
    assert err <= 1e-5, err
