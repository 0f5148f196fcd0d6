"""Recorded PCG iterations (femh_opts.graph): a pair of iterations is
captured into a hipGraph once per context and shape, then replayed until the
load cases converge, with the convergence test on the device.  Every replay
must give what launching the same kernels one by one gives: the same C_H bit
for bit and the same iteration counts, over several solves in a row on one
context (the recorded pair is reused), with a smaller last batch (a second
recording), in FP64 and in mixed precision, with the multigrid and the Jacobi
preconditioner.  Eager solves on another context run between the replays."""
import numpy as np
import pytest

from fem_amd import Homogenizer, homogenize_batch
from fem_amd import workloads as W

pytestmark = pytest.mark.gpu
NU = 0.3


@pytest.mark.parametrize("precision,precond", [("fp64", "mg"), ("mixed", "mg"), ("fp64", "jacobi")])
def test_graph_replays_equal_eager_launches(precision, precond):
    shape = (24, 24, 24)
    rho = np.stack([W.simp_density(shape, 30 + s, 3.0) for s in range(7)])
    kw = dict(nu=NU, tol=1e-10, batch=4, precision=precision, precond=precond, max_iter=20000,
              ch_tol=1e-7 if precision == "mixed" else 0.0)
    hg = Homogenizer(shape, graph=True, **kw)
    he = Homogenizer(shape, graph=False, **kw)
    for rep in range(3):
        Cg, sg = hg.homogenize_many(rho)
        Ce, se = he.homogenize_many(rho)
        assert np.array_equal(Cg, Ce), (rep, np.abs(Cg - Ce).max())
        for a, b in zip(sg, se):
            assert a["iters"] == b["iters"] and a["outer"] == b["outer"], (rep, a, b)
            assert a["relres"] == b["relres"]
        G = np.random.default_rng(rep).standard_normal((3, 6, 6))
        assert np.array_equal(hg.gradient(G, nb=3), he.gradient(G, nb=3))
    hg.close()
    he.close()


def test_graph_uneven_batch_equals_eager():
    grids = W.batch_uneven(20, 9, edges=(8, 16, 24))
    kw = dict(nu=NU, tol=1e-10, batch=4, precision="mixed", ch_tol=1e-7, max_iter=20000)
    Cg, _ = homogenize_batch(grids, graph=True, **kw)
    Ce, _ = homogenize_batch(grids, graph=False, **kw)
    assert np.array_equal(Cg, Ce), np.abs(Cg - Ce).max()
