"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle."""
import numpy as np
import pytest

from fem_amd import Homogenizer
from fem_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NU = 0.3


def _rand_field(shape, seed):
    return np.random.default_rng(seed).standard_normal((3,) + (int(np.prod(shape)),))


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 4, 5), (8, 8, 8), (6, 4, 10)])
def test_fine_apply_matches_oracle(shape):
    E = np.random.default_rng(1).uniform(1e-3, 1.0, shape)
    h = Homogenizer(shape, nu=NU, precond="jacobi")
    h.prepare(E)
    u = _rand_field(shape, 2)
    y = h.apply_level(0, u).reshape(3, -1)
    y_ref = O.apply(E, NU, u)
    assert np.abs(y - y_ref).max() <= 1e-12 * np.abs(y_ref).max()


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 8, 4), (32, 32, 32), (32, 48, 40)])
def test_coarse_operators_match_oracle_galerkin(shape):
    """Every coarse level's operator against the oracle's Galerkin hierarchy;
    levels with 16x16 planes or more run the LDS-tiled kernel (full and
    partial tiles)."""
    E = W.binary_modulus(W.binary_cube(shape, 3))
    h = Homogenizer(shape, nu=NU, precond="mg")
    h.prepare(E)
    mg = O.MG(E, NU)
    for l in range(1, mg.L):
        d = h.level_dims(l)
        assert d == mg.dims[l]
        u = _rand_field(d, 10 + l).reshape(-1)
        y = h.apply_level(l, u)
        y_ref = mg.apply(l, u)
        assert np.abs(y - y_ref).max() <= 1e-11 * np.abs(y_ref).max(), l


@pytest.mark.parametrize("precond", ["jacobi", "mg"])
def test_homogenize_fp64_matches_oracle_8(precond):
    E = W.binary_modulus(W.binary_cube(8, 5))
    ref = O.homogenize(E, NU, precond="jacobi", tol=1e-12)
    h = Homogenizer((8, 8, 8), nu=NU, precond=precond, tol=1e-10)
    C, st = h.homogenize_modulus(E)
    err = np.linalg.norm(C - ref["C"]) / np.linalg.norm(ref["C"])
    assert err <= 1e-10, (err, st)


def test_homogeneous_cell_is_exact():
    lam, mu = O.lame(1.0, NU)
    Ciso = np.zeros((6, 6))
    Ciso[:3, :3] = lam
    Ciso[range(3), range(3)] += 2 * mu
    Ciso[3, 3] = Ciso[4, 4] = Ciso[5, 5] = mu
    h = Homogenizer((8, 8, 8), nu=NU)
    C, st = h.homogenize_modulus(np.ones((8, 8, 8)))
    assert np.abs(C - Ciso).max() <= 1e-14


TILES = {1: (16, 16), 2: (8, 32), 3: (16, 32)}


@pytest.mark.parametrize("tile", [1, 2, 3])
@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 4, 5), (6, 4, 10), (20, 33, 40), (9, 40, 64)])
def test_fine_apply_every_tile_shape_matches_oracle(tile, shape):
    """The fine operator with each tile of owned node columns (TJ x TK), with
    the built-in and an uneven plane count per workgroup: tiles wider than
    the grid, partial tiles and periodic wrap on every side."""
    E = np.random.default_rng(7).uniform(1e-3, 1.0, shape)
    u = _rand_field(shape, 8)
    y_ref = O.apply(E, NU, u)
    for planes in (0, 3):
        h = Homogenizer(shape, nu=NU, precond="jacobi")
        h.set_launch_shapes([(tile << 8) | planes])
        h.prepare(E)
        y = h.apply_level(0, u).reshape(3, -1)
        assert h.launch_shapes()[0] == (tile << 8) | planes
        assert np.abs(y - y_ref).max() <= 1e-12 * np.abs(y_ref).max(), (tile, planes)
        h.close()


@pytest.mark.parametrize("tile", [2, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_with_each_tile_shape_matches_oracle(tile, precision):
    """A full multigrid solve (fused dot products of 256- and 512-thread
    workgroups) with a non-default tile, against the oracle."""
    shape = (24, 20, 36)
    rho = W.simp_density(shape, 11, 3.0)
    E = W.simp_modulus(rho)
    ref = O.homogenize(E, NU, precond="mg", tol=1e-12)
    h = Homogenizer(shape, nu=NU, precision=precision, tol=1e-11, ch_tol=0.0, batch=2)
    h.set_launch_shapes([tile << 8])
    C, st = h.homogenize_many(np.stack([rho, rho]))
    for s in range(2):
        err = np.linalg.norm(C[s] - ref["C"]) / np.linalg.norm(ref["C"])
        assert err <= 1e-9, (tile, precision, err, st[s])
