"""GPU tests of the full homogenization path through the C ABI: golden C_H of
the seeded workloads, mixed precision, adjoint, batch, error handling."""
import json
import os

import numpy as np
import pytest

from fem_amd import FemhError, Homogenizer, homogenize_batch
from fem_amd import workloads as W
from fem_amd.femh import BAD_ARG, NO_FIELDS, NONFINITE
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ch_goldens.json")))
NU = 0.3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a).ravel() - np.asarray(b).ravel()) / np.linalg.norm(b))


@pytest.mark.parametrize("seed", [0, 1])
def test_binary16_jacobi_matches_golden(seed):
    g = GOLD[f"binary16_s{seed}"]
    E = W.binary_modulus(W.binary_cube(16, seed), 1.0, g["E_void"])
    h = Homogenizer((16, 16, 16), nu=NU, precond="jacobi", tol=1e-10, max_iter=20000)
    C, st = h.homogenize_modulus(E)
    assert rel(C, g["C"]) <= 1e-10, st


@pytest.mark.parametrize("kernel", ["tiled", "dense"])
def test_simp64_multigrid_fp64_matches_golden(kernel):
    g = GOLD["simp64_s0"]
    rho = W.simp_density(64, 0, 4.0)
    h = Homogenizer((64, 64, 64), nu=NU, penal=3.0, tol=1e-10, apply_kernel=kernel)
    C, st = h.homogenize(rho)
    assert rel(C, g["C"]) <= 1e-10, st
    assert max(st["iters"]) <= 40


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_simp64_three_level_vcycle_matches_golden(precision):
    """64^3 SIMP cell with a 3-level hierarchy (64, 32, 16): the 16^3
    coarsest level is too large for the dense inverse and is solved by a
    Chebyshev polynomial; V(2,2)-PCG to 1e-10."""
    g = GOLD["simp64_s0"]
    rho = W.simp_density(64, 0, 4.0)
    h = Homogenizer((64, 64, 64), nu=NU, penal=3.0, tol=1e-10, levels=3, nu1=2, nu2=2,
                    precision=precision, ch_tol=0.0)
    C, st = h.homogenize(rho)
    assert st["levels"] == 3
    assert rel(C, g["C"]) <= 1e-10, st
    assert max(st["relres"]) <= 1e-10, st
    if precision == "fp64":
        assert max(st["iters"]) <= 200
    # (mixed, ch_tol 0: the FP32 corrections of this 1e9-contrast cell stall
    # near 1e-10, so a load case may be finished by FP64 PCG; only the result
    # is checked)


def test_tiled_and_dense_apply_agree_on_odd_grid():
    shape = (5, 7, 9)
    E = np.random.default_rng(3).uniform(1e-3, 1.0, shape)
    u = np.random.default_rng(4).standard_normal(3 * E.size)
    y = []
    for kernel in ("tiled", "dense"):
        h = Homogenizer(shape, nu=NU, precond="jacobi", apply_kernel=kernel)
        h.prepare(E)
        y.append(h.apply_level(0, u))
    assert np.abs(y[0] - y[1]).max() <= 1e-13 * np.abs(y[1]).max()


@pytest.mark.parametrize("shape", [(24, 24, 24), (40, 16, 32), (33, 18, 17)])
def test_tiled_apply_matches_oracle_on_partial_tiles(shape):
    E = np.random.default_rng(5).uniform(1e-3, 1.0, shape)
    u = np.random.default_rng(6).standard_normal(3 * E.size)
    h = Homogenizer(shape, nu=NU, precond="jacobi")
    h.prepare(E)
    y = h.apply_level(0, u)
    yr = O.apply(E, NU, u).ravel()
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()


def test_mixed_precision_meets_1e6_against_fp64_golden():
    g = GOLD["simp64_s0"]
    rho = W.simp_density(64, 0, 4.0)
    h = Homogenizer((64, 64, 64), nu=NU, precision="mixed", tol=1e-10, ch_tol=1e-6)
    C, st = h.homogenize(rho)
    err = rel(C, g["C"])
    assert err <= 1e-6, st
    # the a-posteriori estimate bounds the measured error
    assert err <= st["ch_err_est"] <= 1e-6, (err, st["ch_err_est"])


def test_mixed_to_tight_tolerance_matches_fp64():
    g = GOLD["binary16_s1"]
    E = W.binary_modulus(W.binary_cube(16, 1), 1.0, g["E_void"])
    h = Homogenizer((16, 16, 16), nu=NU, precision="mixed", tol=1e-10)
    C, st = h.homogenize_modulus(E)
    assert rel(C, g["C"]) <= 1e-9, st
    assert max(st["relres"]) <= 1e-10


def test_fields_zero_mean_and_solve_the_system():
    E = W.simp_modulus(W.simp_density(16, 0, 4.0))
    h = Homogenizer((16, 16, 16), nu=NU, tol=1e-10)
    C, _ = h.homogenize_modulus(E)
    u = h.fields()
    assert np.abs(u.mean(axis=2)).max() < 1e-12
    for lc in range(6):
        r = O.rhs(E, NU, lc) - O.apply(E, NU, u[lc])
        assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(O.rhs(E, NU, lc))
    assert rel(O.ch_energy(E, NU, u), C) < 1e-13


def test_gradient_matches_oracle_energy_terms():
    rho = W.simp_density(8, 3, 2.0)
    h = Homogenizer((8, 8, 8), nu=NU, penal=3.0, tol=1e-12)
    C, _ = h.homogenize(rho)
    Wg = h.energy_terms()
    E = W.simp_modulus(rho)
    ref = O.homogenize(E, NU, precond="mg", tol=1e-13)
    Wr = O.energy_terms(rho.shape, NU, ref["u"])
    assert np.abs(Wg - Wr).max() <= 1e-9 * np.abs(Wr).max()
    G = np.random.default_rng(8).standard_normal((6, 6))
    g = h.gradient(G)
    gr = O.grad_E(rho.shape, NU, ref["u"], G) * 3.0 * rho ** 2
    assert np.abs(g - gr).max() <= 1e-8 * np.abs(gr).max()


def test_gradient_central_differences_64_voxels():
    """dJ/drho for J = ||C - C_t||_F^2 against central differences (h = 1e-6)."""
    n = 8
    rho = W.simp_density(n, 4, 2.0)
    h = Homogenizer((n, n, n), nu=NU, penal=3.0, tol=1e-13, max_iter=20000)
    lam, mu = O.lame(0.3, 0.25)
    Ct = np.zeros((6, 6))
    Ct[:3, :3] = lam
    Ct[range(3), range(3)] += 2 * mu
    Ct[3, 3] = Ct[4, 4] = Ct[5, 5] = mu

    def J(r):
        C, _ = h.homogenize(r)
        return float(np.sum((C - Ct) ** 2)), C

    J0, C0 = J(rho)
    g = h.gradient(2.0 * (C0 - Ct))
    idx = W.uniform(11, 64) * rho.size
    worst = 0.0
    for f in idx.astype(int):
        e = np.unravel_index(f, rho.shape)
        hstep = 1e-6
        rp, rm = rho.copy(), rho.copy()
        rp[e] += hstep
        rm[e] -= hstep
        fd = (J(rp)[0] - J(rm)[0]) / (2 * hstep)
        worst = max(worst, abs(fd - g[e]) / np.abs(g).max())
    assert worst <= 1e-5, worst


def test_gradient_euler_identity_for_pure_power_law():
    # E = rho^p (Emin = 0): sum_e rho_e dC/drho_e = p C
    rho = W.simp_density(16, 6, 4.0)
    h = Homogenizer((16, 16, 16), nu=NU, penal=3.0, Emin=0.0, tol=1e-12)
    C, _ = h.homogenize(rho)
    S = np.zeros((6, 6))
    for a in range(6):
        for b in range(a, 6):
            G = np.zeros((6, 6))
            G[a, b] = 1.0
            S[a, b] = S[b, a] = float(np.sum(rho * h.gradient(G)))
    assert np.abs(S - 3.0 * C).max() <= 1e-9 * np.abs(C).max()


def test_autograd_gradcheck():
    torch = pytest.importorskip("torch")
    from fem_amd.autograd import homogenize
    rho = torch.from_numpy(W.simp_density(4, 9, 1.5)).requires_grad_(True)
    h = Homogenizer((4, 4, 4), nu=NU, penal=3.0, tol=1e-14, max_iter=20000)
    assert torch.autograd.gradcheck(lambda r: homogenize(r, h), (rho,), eps=1e-6, atol=1e-7,
                                    rtol=1e-4)


def test_batch_matches_individual_solves():
    grids = W.batch_uneven(6, 3, edges=(8, 12, 16))
    C, st = homogenize_batch(grids, nu=NU, tol=1e-10)
    for g, c in zip(grids, C):
        h = Homogenizer(g.shape, nu=NU, tol=1e-10)
        ci, _ = h.homogenize_modulus(g)
        assert rel(c, ci) <= 1e-12


def test_homogeneous_and_laminate_exact_on_gpu():
    from tests.test_oracle_pins import c_iso, laminate_closed_form
    h = Homogenizer((8, 8, 8), nu=0.45)
    C, _ = h.homogenize_modulus(3.0 * np.ones((8, 8, 8)))
    assert np.abs(C - c_iso(3.0, 0.45)).max() <= 1e-13
    Es = np.array([1.0, 0.05, 0.3, 0.3, 1.0, 1.0, 0.01, 0.5])
    E = np.broadcast_to(Es[:, None, None], (8, 8, 8)).copy()
    h = Homogenizer((8, 8, 8), nu=NU, tol=1e-12)
    C, _ = h.homogenize_modulus(E)
    Cr = laminate_closed_form(Es, np.full(8, 1 / 8), NU)
    assert np.abs(C - Cr).max() <= 1e-10 * np.abs(Cr).max()


def test_error_paths():
    with pytest.raises(FemhError) as e:
        Homogenizer((1, 4, 4))
    assert e.value.status == BAD_ARG
    with pytest.raises(FemhError) as e:
        Homogenizer((4, 4, 4), nu=0.5)
    assert e.value.status == BAD_ARG
    h = Homogenizer((4, 4, 4))
    with pytest.raises(FemhError) as e:
        h.gradient(np.eye(6))
    assert e.value.status == NO_FIELDS
    bad = np.full((4, 4, 4), 0.5)
    bad[1, 2, 3] = np.nan
    with pytest.raises(FemhError) as e:
        h.homogenize(bad)
    assert e.value.status == NONFINITE
    with pytest.raises(FemhError) as e:
        h.homogenize_modulus(np.zeros((4, 4, 4)))
    assert e.value.status == BAD_ARG
    # the context stays usable after errors
    C, _ = h.homogenize_modulus(np.ones((4, 4, 4)))
    assert np.isfinite(C).all()


def test_not_converged_is_reported():
    E = W.binary_modulus(W.binary_cube(16, 0))
    h = Homogenizer((16, 16, 16), precond="jacobi", max_iter=3)
    C, st = h.homogenize_modulus(E, check=False)
    assert max(st["iters"]) == 3
    with pytest.raises(FemhError):
        h.homogenize_modulus(E)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_warm_start_reuses_previous_fields(precision):
    """An optimisation loop re-solves a slightly changed cell: warm starts
    from the previous fields need fewer iterations and give the cold result."""
    rho = W.simp_density(32, 3, 3.0)
    rho2 = np.clip(rho + 0.01 * np.sin(np.arange(rho.size)).reshape(rho.shape), 1e-3, 1.0)
    kw = dict(nu=NU, tol=1e-10, precision=precision, ch_tol=1e-9 if precision == "mixed" else 0.0)
    cold = Homogenizer(rho.shape, **kw)
    C1c, _ = cold.homogenize(rho)
    C2c, st2c = cold.homogenize(rho2)
    warm = Homogenizer(rho.shape, warm_start=True, **kw)
    C1, st1 = warm.homogenize(rho)
    assert rel(C1, C1c) <= 1e-12
    # the same cell again: already converged, no iterations
    C1b, st1b = warm.homogenize(rho)
    assert max(st1b["iters"]) == 0, st1b
    assert rel(C1b, C1c) <= (1e-12 if precision == "fp64" else 1e-8)
    C2, st2 = warm.homogenize(rho2)
    assert sum(st2["iters"]) < sum(st2c["iters"]), (st2["iters"], st2c["iters"])
    assert rel(C2, C2c) <= (1e-10 if precision == "fp64" else 1e-7)
