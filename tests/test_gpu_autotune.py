"""GPU tests of the launch-shape autotuner (femh_opts.autotune): tuned shapes
are valid and cached for the process, and solves with them agree with the
built-in shapes (the fine-operator shape only reorders partial sums)."""
import numpy as np
import pytest

from fem_amd import Homogenizer, homogenize_batch
from fem_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_autotuned_fp64_solve_matches_builtin_shapes():
    n = 32
    Es = np.stack([W.simp_modulus(W.simp_density(n, s, 4.0)) for s in (1, 2)])
    ht = Homogenizer((n, n, n), nu=0.3, tol=1e-12, batch=2, autotune=True)
    hb = Homogenizer((n, n, n), nu=0.3, tol=1e-12, batch=2)
    Ct, st = ht.homogenize_many(Es, modulus=True)
    Cb, _ = hb.homogenize_many(Es, modulus=True)
    for s in range(2):
        assert rel(Ct[s], Cb[s]) <= 1e-10
    shapes = ht.launch_shapes()
    assert len(shapes) == st[0]["levels"] >= 3
    # fine operator: bits 0-7 node planes per workgroup, bits 8-15 tile of
    # owned node columns (0 = built-in, 1 = 16 x 16, 2 = 8 x 32, 3 = 16 x 32)
    assert (shapes[0] & 0xFF) in (0, 2, 4, 8, 16, 32) and (shapes[0] >> 8) in (0, 1, 2, 3)
    assert all(c in (0, 1, 2, 4, 8) for c in shapes[1:])
    assert hb.launch_shapes() == [0] * len(shapes)
    # a second context of the same shape takes the cached choice
    h2 = Homogenizer((n, n, n), nu=0.3, tol=1e-12, batch=2, autotune=True)
    h2.homogenize_many(Es, modulus=True)
    assert h2.launch_shapes() == shapes
    # a context pinned to the tuned shapes gives the tuned result exactly
    hp = Homogenizer((n, n, n), nu=0.3, tol=1e-12, batch=2)
    hp.set_launch_shapes(shapes)
    Cp, _ = hp.homogenize_many(Es, modulus=True)
    assert np.array_equal(Cp, Ct)
    for h in (ht, hb, h2, hp):
        h.close()


def test_autotuned_mixed_uneven_batch_within_1e6():
    grids = W.batch_uneven(24, 11)
    C, _ = homogenize_batch(grids, nu=0.3, precision="mixed", tol=1e-10, ch_tol=1e-7, inner_tol=3e-4,
                            max_iter=20000, batch=8, autotune=True)
    for s, g in enumerate(grids):
        h = Homogenizer(g.shape, nu=0.3, tol=1e-11, max_iter=20000, interp="modulus")
        C64, _ = h.homogenize_modulus(g)
        h.close()
        assert rel(C[s], C64) <= 1e-6, (s, g.shape, rel(C[s], C64))
