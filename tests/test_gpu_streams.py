"""GPU test of concurrent batches in a many-call (femh_opts.streams): every
cell's C_H is bit-identical to the one-stream call, and the gradient call
still refers to the last batch."""
import numpy as np
import pytest

from fem_amd import Homogenizer
from fem_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_streams_give_identical_results_and_last_batch_fields(precision):
    n, B = 16, 2
    rhos = np.stack([W.simp_density(n, s, 3.0) for s in range(7)])  # 4 batches, the last one short
    kw = dict(nu=0.3, penal=3.0, precision=precision, tol=1e-10, ch_tol=1e-7, batch=B)
    h1 = Homogenizer((n, n, n), **kw)
    h3 = Homogenizer((n, n, n), streams=3, **kw)
    C1, s1 = h1.homogenize_many(rhos)
    C3, s3 = h3.homogenize_many(rhos)
    assert np.array_equal(C1, C3)
    assert [s["iters"] for s in s1] == [s["iters"] for s in s3]
    G = np.random.default_rng(2).standard_normal((1, 6, 6))
    g1 = h1.gradient(G, nb=1)
    g3 = h3.gradient(G, nb=1)
    assert np.array_equal(g1, g3)
    h1.close()
    h3.close()


def test_resident_batch_object_repeats_exactly():
    """BatchHomogenizer keeps its per-size sub-contexts between calls (several
    per size when batches of one size are in flight together); a repeated
    call and the one-shot helper give the same tensors."""
    from fem_amd import BatchHomogenizer, homogenize_batch
    grids = W.batch_uneven(40, 5, edges=(16, 24))
    kw = dict(nu=0.3, precision="mixed", tol=1e-10, ch_tol=1e-7, inner_tol=3e-4, max_iter=20000, batch=4)
    bh = BatchHomogenizer(**kw)
    C1, _ = bh.run(grids)
    C2, _ = bh.run(grids)
    bh.close()
    C3, _ = homogenize_batch(grids, **kw)
    assert np.array_equal(C1, C2) and np.array_equal(C1, C3)
