"""CPU pins of the oracle (and of the engine's host-side element tables).

Each pin is independent of the oracle's own formulas:
  * tests/golden/ref_1K.txt -- the reference's own 24x24 Ke fixture (tests/1.K,
    E=1, nu=0.45, counter-clockwise node order)
  * lk_H8 -- the closed-form unit hex stiffness of matlab/top3d.m:101-147
    (a different derivation: hand-integrated coefficients)
  * oracle/_ref -- the reference's femlib compiled from its sources
  * brute-force sparse assembly with explicit periodic node identification
  * laminate closed forms (Voigt/Reuss-type averages), homogeneous cell,
    scaling, symmetry, 90-degree rotations
  * adjoint: central finite differences and the Euler identity
"""
import ctypes
import os

import numpy as np
import pytest

from fem_amd import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
NU = 0.3


def c_iso(E, nu):
    lam, mu = O.lame(E, nu)
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def load_1K():
    tok = open(os.path.join(HERE, "golden", "ref_1K.txt")).read().split()
    n, m = int(tok[0]), int(tok[1])
    K = np.zeros((n, m))
    for t in range(2, len(tok), 3):
        K[int(tok[t]), int(tok[t + 1])] = float(tok[t + 2])
    return K


def ccw_perm():
    # 1.K corner order (x,y,z): 000,100,110,010,001,101,111,011 -> local 4x+2y+z
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    P = []
    for (x, y, z) in order:
        v = 4 * x + 2 * y + z
        P += [3 * v, 3 * v + 1, 3 * v + 2]
    return np.array(P)


def lk_h8(nu):
    """Unit-cube trilinear hex stiffness from the hand-integrated coefficient
    table of matlab/top3d.m:101-147 (E = 1), in that file's node order."""
    A = np.array([[32, 6, -8, 6, -6, 4, 3, -6, -10, 3, -3, -3, -4, -8],
                  [-48, 0, 0, -24, 24, 0, 0, 0, 12, -12, 0, 12, 12, 12]], float)
    k = (A.T @ np.array([1.0, nu])) / 144.0
    k = np.concatenate([[0.0], k])  # 1-based like the source
    K1 = [[1, 2, 2, 3, 5, 5], [2, 1, 2, 4, 6, 7], [2, 2, 1, 4, 7, 6],
          [3, 4, 4, 1, 8, 8], [5, 6, 7, 8, 1, 2], [5, 7, 6, 8, 2, 1]]
    K2 = [[9, 8, 12, 6, 4, 7], [8, 9, 12, 5, 3, 5], [10, 10, 13, 7, 4, 6],
          [6, 5, 11, 9, 2, 10], [4, 3, 5, 2, 9, 12], [11, 4, 6, 12, 10, 13]]
    K3 = [[6, 7, 4, 9, 12, 8], [7, 6, 4, 10, 13, 10], [5, 5, 3, 8, 12, 9],
          [9, 10, 2, 6, 11, 5], [12, 13, 10, 11, 6, 4], [2, 12, 9, 4, 5, 3]]
    K4 = [[14, 11, 11, 13, 10, 10], [11, 14, 11, 12, 9, 8], [11, 11, 14, 12, 8, 9],
          [13, 12, 12, 14, 7, 7], [10, 9, 8, 7, 14, 11], [10, 8, 9, 7, 11, 14]]
    K5 = [[1, 2, 8, 3, 5, 4], [2, 1, 8, 4, 6, 11], [8, 8, 1, 5, 11, 6],
          [3, 4, 5, 1, 8, 2], [5, 6, 11, 8, 1, 8], [4, 11, 6, 2, 8, 1]]
    K6 = [[14, 11, 7, 13, 10, 12], [11, 14, 7, 12, 9, 2], [7, 7, 14, 10, 2, 9],
          [13, 12, 10, 14, 7, 11], [10, 9, 2, 7, 14, 7], [12, 2, 9, 11, 7, 14]]
    M = {n: k[np.array(v)] for n, v in
         dict(K1=K1, K2=K2, K3=K3, K4=K4, K5=K5, K6=K6).items()}
    KE = np.block([[M["K1"], M["K2"], M["K3"], M["K4"]],
                   [M["K2"].T, M["K5"], M["K6"], M["K3"].T],
                   [M["K3"].T, M["K6"], M["K5"].T, M["K2"].T],
                   [M["K4"], M["K3"], M["K2"], M["K1"].T]])
    return KE / ((nu + 1) * (1 - 2 * nu))


def top3d_perm():
    """top3d's hex node order is counter-clockwise in x-y with y pointing down
    the element (nodes (0,1,0),(1,1,0),(1,0,0),(0,0,0) then z=1), found by
    matching the rigid-body null space; returned as a map to local 4x+2y+z."""
    order = [(0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]
    P = []
    for (x, y, z) in order:
        v = 4 * x + 2 * y + z
        P += [3 * v, 3 * v + 1, 3 * v + 2]
    return np.array(P)


# ---------------------------------------------------------------- element
def test_oracle_ke_matches_reference_1K_fixture():
    K1 = load_1K()
    K = O.ke_hex(1.0, 0.45)
    P = ccw_perm()
    assert np.abs(K[np.ix_(P, P)] - K1).max() < 1e-13


def test_oracle_ke_matches_top3d_closed_form():
    P = top3d_perm()
    for nu in (0.3, 0.45, 0.0, -0.2):
        K = O.ke_hex(1.0, nu)
        KE = lk_h8(nu)
        # top3d's y axis points the other way: flip the sign of y dofs
        S = np.ones(24)
        S[1::3] = -1
        Kp = K[np.ix_(P, P)] * np.outer(S, S)
        assert np.abs(Kp - KE).max() < 1e-13 * np.abs(KE).max(), nu


def test_engine_tables_match_oracle_ke():
    from fem_amd.femh import element_stiffness
    for nu in (0.3, 0.45, 0.1):
        assert np.abs(element_stiffness(2.5, nu) - O.ke_hex(2.5, nu)).max() < 1e-14


def _ref_lib(kind):
    path = os.path.join(os.path.dirname(HERE), "oracle", "_ref", f"libref_femlib_{kind}.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    return ctypes.CDLL(path)


def test_oracle_ke_matches_compiled_reference_femlib():
    L = _ref_lib("double")
    assert L.ref_scalar_bytes() == 8
    out = np.zeros(576)
    for nu in (0.3, 0.45):
        lam, mu = O.lame(1.0, nu)
        assert L.ref_ke_hex(ctypes.c_double(mu), ctypes.c_double(lam),
                            out.ctypes.data_as(ctypes.c_void_p)) == 0
        Kr = out.reshape(24, 24)
        # the reference rounds the Gauss abscissa to float (femlib/src/Quadrature.cpp:5-6)
        Kf = O.ke_hex(1.0, nu, float_gauss=True)
        assert np.abs(Kr - Kf).max() < 1e-15 * 10
        rel = np.abs(Kr - O.ke_hex(1.0, nu)).max() / np.abs(Kr).max()
        assert 1e-9 < rel < 1e-7  # the documented ~1.8e-8 abscissa quirk


def test_oracle_ke_matches_compiled_reference_float_build():
    L = _ref_lib("float")
    assert L.ref_scalar_bytes() == 4
    out = np.zeros(576)
    lam, mu = O.lame(1.0, 0.3)
    L.ref_ke_hex(ctypes.c_double(mu), ctypes.c_double(lam), out.ctypes.data_as(ctypes.c_void_p))
    K = O.ke_hex(1.0, 0.3)
    assert np.abs(out.reshape(24, 24) - K).max() / np.abs(K).max() < 1e-6


def test_ke_invariants():
    K = O.ke_hex(1.0, NU)
    assert np.abs(K - K.T).max() < 1e-15
    w = np.linalg.eigvalsh(K)
    assert (w > -1e-13).all()
    assert (np.abs(w) < 1e-12).sum() == 6  # rigid motions
    assert np.abs(K.reshape(24, 8, 3).sum(axis=1)).max() < 1e-14  # translations


# ---------------------------------------------------------------- operator
def assembled_periodic(E, nu):
    """Explicit sparse assembly: every voxel's Ke added at its wrapped nodes."""
    import scipy.sparse as sp
    nx, ny, nz = E.shape
    V = E.size
    K = O.ke_hex(1.0, nu)
    rows, cols, vals = [], [], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                nodes = []
                for v in range(8):
                    bi, bj, bk = (v >> 2) & 1, (v >> 1) & 1, v & 1
                    nodes.append((((i + bi) % nx) * ny + (j + bj) % ny) * nz + (k + bk) % nz)
                dof = [c * V + n for n in nodes for c in range(3)]
                for a in range(24):
                    for b in range(24):
                        rows.append(dof[a])
                        cols.append(dof[b])
                        vals.append(E[i, j, k] * K[a, b])
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * V, 3 * V))


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 4), (4, 4, 4), (2, 3, 2)])
def test_oracle_apply_matches_brute_force_assembly(shape):
    rng = np.random.default_rng(0)
    E = rng.uniform(0.01, 1.0, shape)
    A = assembled_periodic(E, NU)
    u = rng.standard_normal(3 * E.size)
    assert np.abs(O.apply(E, NU, u).ravel() - A @ u).max() < 1e-12
    assert np.abs(O.diag(E, NU).ravel() - A.diagonal()).max() < 1e-14


def test_oracle_operator_symmetric_and_singular_on_translations():
    rng = np.random.default_rng(1)
    E = rng.uniform(0.01, 1.0, (4, 3, 5))
    u, v = rng.standard_normal((2, 3 * E.size))
    Au, Av = O.apply(E, NU, u).ravel(), O.apply(E, NU, v).ravel()
    assert abs(u @ Av - v @ Au) < 1e-12 * abs(u @ Av)
    t = np.zeros((3, E.size))
    t[1] = 1.0
    assert np.abs(O.apply(E, NU, t)).max() < 1e-13


def test_oracle_rhs_is_minus_K_times_affine_field():
    # b(lc) = -K u_aff for the continuous affine field restricted to nodes,
    # which is periodic-consistent only element-locally: check on the
    # assembled matrix with the affine field written per element
    rng = np.random.default_rng(2)
    E = rng.uniform(0.1, 1.0, (3, 3, 3))
    K = O.ke_hex(1.0, NU)
    for lc in range(6):
        eps = np.zeros((3, 3))
        d = {0: (0, 0), 1: (1, 1), 2: (2, 2), 3: (1, 2), 4: (0, 2), 5: (0, 1)}[lc]
        eps[d] = 1.0 if lc < 3 else 0.5
        eps[d[::-1]] = eps[d]
        b = np.zeros((3, E.size))
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    ue = np.zeros(24)
                    for v in range(8):
                        x = np.array([(v >> 2) & 1, (v >> 1) & 1, v & 1], float)
                        ue[3 * v:3 * v + 3] = eps @ x
                    f = E[i, j, k] * (K @ ue)
                    for v in range(8):
                        n = (((i + ((v >> 2) & 1)) % 3) * 3 + (j + ((v >> 1) & 1)) % 3) * 3 + (k + (v & 1)) % 3
                        b[:, n] -= f[3 * v:3 * v + 3]
        assert np.abs(O.rhs(E, NU, lc) - b).max() < 1e-14


def test_oracle_multigrid_coarse_operator_is_galerkin_product():
    import scipy.sparse as sp
    E = W.binary_modulus(W.binary_cube((4, 4, 4), 5), 1.0, 1e-3)
    mg = O.MG(E, NU, levels=2)
    A = assembled_periodic(E, NU)
    Vf, Vc = 64, 8
    # explicit trilinear periodic prolongation, column by column
    P = np.zeros((3 * Vf, 3 * Vc))
    for col in range(3 * Vc):
        e = np.zeros(3 * Vc)
        e[col] = 1.0
        P[:, col] = mg.prolong_add(0, e, np.zeros(3 * Vf))
    Ac = P.T @ (A @ P)
    for col in range(3 * Vc):
        e = np.zeros(3 * Vc)
        e[col] = 1.0
        assert np.abs(mg.apply(1, e) - Ac[:, col]).max() < 1e-13
    r = np.random.default_rng(3).standard_normal(3 * Vf)
    assert np.abs(mg.restrict(0, r) - P.T @ r).max() < 1e-13


# ---------------------------------------------------------------- C_H
def test_oracle_homogeneous_cell_gives_isotropic_tensor():
    for nu in (0.3, 0.45):
        r = O.homogenize(2.0 * np.ones((4, 4, 4)), nu)
        assert np.abs(r["C"] - c_iso(2.0, nu)).max() < 1e-14


def laminate_closed_form(Es, fr, nu):
    """Layers normal to x with moduli Es (volume fractions fr)."""
    lam = np.array([O.lame(E, nu)[0] for E in Es])
    mu = np.array([O.lame(E, nu)[1] for E in Es])
    M = lam + 2 * mu
    av = lambda a: float(np.sum(fr * a))  # noqa: E731
    h = 1.0 / av(1.0 / M)
    C = np.zeros((6, 6))
    C[0, 0] = h
    C[0, 1] = C[0, 2] = C[1, 0] = C[2, 0] = av(lam / M) * h
    C[1, 1] = C[2, 2] = av(M - lam ** 2 / M) + av(lam / M) ** 2 * h
    C[1, 2] = C[2, 1] = av(lam - lam ** 2 / M) + av(lam / M) ** 2 * h
    C[3, 3] = av(mu)
    C[4, 4] = C[5, 5] = 1.0 / av(1.0 / mu)
    return C


@pytest.mark.parametrize("precond", ["jacobi", "mg"])
def test_oracle_laminate_matches_closed_form(precond):
    Es = np.array([1.0, 0.05, 0.3, 0.3, 1.0, 1.0, 0.01, 0.5])
    E = np.broadcast_to(Es[:, None, None], (8, 4, 4)).copy()
    r = O.homogenize(E, NU, precond=precond, tol=1e-13, max_iter=20000)
    Cref = laminate_closed_form(Es, np.full(8, 1 / 8), NU)
    assert np.abs(r["C"] - Cref).max() < 1e-11 * np.abs(Cref).max()


def test_oracle_ch_scaling_symmetry_rotation_and_bounds():
    E = W.binary_modulus(W.binary_cube(6, 11), 1.0, 1e-2)
    C = O.homogenize(E, NU, precond="jacobi", tol=1e-13, max_iter=20000)["C"]
    C2 = O.homogenize(7.0 * E, NU, precond="jacobi", tol=1e-13, max_iter=20000)["C"]
    assert np.abs(C2 - 7.0 * C).max() < 1e-10 * np.abs(C).max()
    assert np.abs(C - C.T).max() < 1e-14
    assert np.linalg.eigvalsh(C).min() > 0
    # swapping the x and y axes of the cell permutes Voigt indices (0 1)(4 3)
    Cs = O.homogenize(np.ascontiguousarray(np.swapaxes(E, 0, 1)), NU, precond="jacobi",
                      tol=1e-13, max_iter=20000)["C"]
    p = [1, 0, 2, 4, 3, 5]
    assert np.abs(Cs - C[np.ix_(p, p)]).max() < 1e-10 * np.abs(C).max()
    # Voigt (arithmetic) and Reuss (harmonic) bounds on the quadratic forms
    Cv = c_iso(E.mean(), NU)
    Cr = np.linalg.inv(np.mean([np.linalg.inv(c_iso(e, NU)) for e in E.ravel()], axis=0))
    assert np.linalg.eigvalsh(Cv - C).min() > -1e-12
    assert np.linalg.eigvalsh(C - Cr).min() > -1e-12


def test_cubic_cell_has_cubic_tensor():
    """A cell invariant under the cubic group has C_H of cubic form: the
    property the engine's cubic-structure mode (two solves, axis swaps and a
    projection onto c11, c12, c44) relies on."""
    E = W.cubic_average(W.binary_modulus(W.binary_cube(6, 5), 1.0, 1e-2))
    assert np.abs(E - np.swapaxes(E, 0, 1)).max() < 1e-14 and np.abs(E - E[::-1]).max() < 1e-14
    C = O.homogenize(E, NU, precond="jacobi", tol=1e-13, max_iter=20000)["C"]
    c11, c12, c44 = C[0, 0], C[0, 1], C[3, 3]
    Cc = np.zeros((6, 6))
    Cc[:3, :3] = c12
    Cc[range(3), range(3)] = c11
    Cc[range(3, 6), range(3, 6)] = c44
    assert np.abs(C - Cc).max() < 1e-10 * c11
    assert abs(c11 - c12 - 2 * c44) > 1e-3 * c11  # cubic, not isotropic


def levin(C, E1, E2, phi, nu):
    """Levin's exact relation for two isotropic phases of equal Poisson ratio,
    expansion 1 in phase 1 and 0 in phase 2: alpha_H = phi I + (S_H I -
    <1/3K> I) / (1/3K1 - 1/3K2), for any microstructure (a uniform pressure
    with a matched temperature strains both phases alike)."""
    k = lambda e: e / (3 * (1 - 2 * nu))
    I = np.array([1, 1, 1, 0, 0, 0.0])
    mean_c = phi / (3 * k(E1)) + (1 - phi) / (3 * k(E2))
    return phi * I + (np.linalg.solve(C, I) - mean_c * I) / (1 / (3 * k(E1)) - 1 / (3 * k(E2)))


def test_oracle_thermal_stress_uniform_expansion_and_levin():
    # a uniform expansion coefficient expands any cell uniformly
    E = W.simp_modulus(W.simp_density(6, 2, 2.0))
    r = O.homogenize(E, NU, precond="jacobi", tol=1e-13, max_iter=20000)
    aH = np.linalg.solve(r["C"], O.thermal_stress(E, NU, np.full(E.shape, 0.7), r["u"]))
    assert np.abs(aH - 0.7 * np.array([1, 1, 1, 0, 0, 0])).max() < 1e-12
    # two phases, anisotropic cell: every component of Levin's relation
    s = W.binary_cube(6, 4) > 0.5
    E = np.where(s, 1.0, 0.05)
    r = O.homogenize(E, NU, precond="jacobi", tol=1e-13, max_iter=20000)
    aH = np.linalg.solve(r["C"], O.thermal_stress(E, NU, s.astype(float), r["u"]))
    ref = levin(r["C"], 1.0, 0.05, s.mean(), NU)
    assert np.abs(aH - ref).max() < 1e-11
    assert np.abs(aH[3:]).max() > 1e-3  # the shear couplings are exercised


def test_oracle_jacobi_and_multigrid_agree():
    E = W.simp_modulus(W.simp_density(16, 3, 4.0))
    a = O.homogenize(E, NU, precond="jacobi", tol=1e-12, max_iter=20000)
    b = O.homogenize(E, NU, precond="mg", tol=1e-12)
    assert np.linalg.norm(a["C"] - b["C"]) / np.linalg.norm(a["C"]) < 1e-11
    assert max(b["iters"]) < max(a["iters"])


def test_goldens_reproduce_with_oracle():
    import json
    G = json.load(open(os.path.join(HERE, "golden", "ch_goldens.json")))
    g = G["binary16_s0"]
    E = W.binary_modulus(W.binary_cube(16, g["seed"]), 1.0, g["E_void"])
    r = O.homogenize(E, g["nu"], precond="jacobi", tol=1e-12, max_iter=20000)
    assert np.linalg.norm(r["C"].ravel() - g["C"]) / np.linalg.norm(g["C"]) < 1e-12


def test_uneven_batch_goldens_regenerate_and_agree_across_solvers():
    """The uneven-batch goldens (written with multigrid-PCG) are the cells the
    generator makes today, and Jacobi-PCG -- a different preconditioner and
    iteration path -- gives the same C_H."""
    import json
    G = json.load(open(os.path.join(HERE, "golden", "ch_goldens.json")))
    keys = sorted(k for k in G if k.startswith("uneven512_c"))
    assert len(keys) == 4
    grids = W.batch_uneven(512, 3)
    for k in keys[:2]:
        g = G[k]
        E = grids[g["cell"]]
        assert E.shape == (16, 16, 16) and abs(E.mean() - g["fill"]) < 1e-15
        r = O.homogenize(E, g["nu"], precond="jacobi", tol=1e-12, max_iter=40000)
        assert r["status"] == 0
        assert np.linalg.norm(r["C"].ravel() - g["C"]) / np.linalg.norm(g["C"]) < 1e-10, k


# ---------------------------------------------------------------- adjoint
def test_oracle_gradient_matches_central_differences():
    # single thread: the reduction order (hence the last bits of J) must not
    # change between the +h and -h solves
    O.set_threads(1)
    rho = W.simp_density(4, 2, 1.5)
    p = 3.0
    G = np.random.default_rng(4).standard_normal((6, 6))
    def J(r):
        return float(np.sum(G * O.homogenize(W.simp_modulus(r, p), NU, precond="jacobi",
                                              tol=1e-14, max_iter=20000)["C"]))
    base = O.homogenize(W.simp_modulus(rho, p), NU, precond="jacobi", tol=1e-14, max_iter=20000)
    g = O.grad_E(rho.shape, NU, base["u"], G) * p * rho ** (p - 1)
    h = 1e-5
    try:
        for idx in [(0, 0, 0), (1, 2, 3), (3, 3, 1), (2, 0, 2)]:
            rp, rm = rho.copy(), rho.copy()
            rp[idx] += h
            rm[idx] -= h
            fd = (J(rp) - J(rm)) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * np.abs(g).max(), (idx, fd, g[idx])
    finally:
        O.set_threads(os.cpu_count() or 1)


def test_oracle_gradient_euler_identity():
    # C(E) is homogeneous of degree 1 in E: sum_e E_e dC/dE_e = C
    E = W.simp_modulus(W.simp_density(6, 5, 2.0))
    r = O.homogenize(E, NU, precond="jacobi", tol=1e-13, max_iter=20000)
    W21 = O.energy_terms(E.shape, NU, r["u"])
    S = (E.ravel()[:, None] * W21).sum(axis=0)
    C = np.zeros((6, 6))
    p = 0
    for a in range(6):
        for b in range(a, 6):
            C[a, b] = C[b, a] = S[p]
            p += 1
    assert np.abs(C - r["C"]).max() < 1e-11 * np.abs(r["C"]).max()


# ---------------------------------------------------------------- inputs
def test_workload_generator_is_deterministic_and_plausible():
    a = W.binary_cube(16, 3)
    b = W.binary_cube(16, 3)
    assert (a == b).all() and abs(a.mean() - 0.5) < 0.03
    g = W.grf(32, 1, 4.0)
    assert abs(g.mean()) < 1e-12 and abs(g.std() - 1) < 1e-12
    u = W.uniform(9, 100000)
    assert abs(u.mean() - 0.5) < 0.01 and u.min() >= 0 and u.max() < 1
    rho = W.simp_density(16, 0, 4.0)
    assert rho.min() >= 1e-3 and rho.max() <= 1.0
    batch = W.batch_uneven(20, 7)
    assert all(x.shape[0] in W.EDGES for x in batch)
