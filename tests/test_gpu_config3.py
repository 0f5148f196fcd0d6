"""The 512-cell uneven batch at its full size: binary cells with edges drawn
from {16, 24, 32, 40, 48}, fill fraction U[0.2, 0.8], E_solid = 1, E_void =
1e-9, nu = 0.3 (fem_amd.workloads.batch_uneven(512, 3)), solved by
femh_homogenize_batch in mixed precision with the settings of the bench's
uneven512 line.  Every cell must lie within 1e-6 (relative Frobenius norm) of
the FP64 path, and the cells with oracle goldens (tests/golden/ch_goldens.json,
written by the CPU oracle) within 1e-6 of the golden in mixed precision and
1e-10 in FP64.  This replaces the reference's OpenMP sample loop
(SamplesGenerator/SamplesGeneratorImpl.cpp:3707-3731)."""
import json
import os

import numpy as np
import pytest

from fem_amd import BatchHomogenizer
from fem_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
NU = 0.3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a).ravel() - np.asarray(b).ravel()) / np.linalg.norm(b))


def test_uneven512_mixed_every_cell_within_1e6_of_fp64_and_goldens():
    grids = W.batch_uneven(512, 3)
    assert sorted({g.shape[0] for g in grids}) == list(W.EDGES)
    kw = dict(nu=NU, tol=1e-10, max_iter=20000, batch=16)
    bm = BatchHomogenizer(precision="mixed", ch_tol=1e-6, inner_tol=3e-4, **kw)
    Cm, _ = bm.run(grids)
    bm.close()
    b64 = BatchHomogenizer(**kw)
    C64, _ = b64.run(grids)
    b64.close()
    errs = np.array([rel(a, b) for a, b in zip(Cm, C64)])
    worst = int(errs.argmax())
    print(f"512 cells: worst mixed-vs-FP64 error {errs[worst]:.2e} (cell {worst}, "
          f"edge {grids[worst].shape[0]}), median {np.median(errs):.2e}")
    assert errs.max() <= 1e-6, (worst, errs[worst])
    G = json.load(open(os.path.join(HERE, "golden", "ch_goldens.json")))
    keys = [k for k in G if k.startswith("uneven512_c")]
    assert keys
    for k in keys:
        g = G[k]
        i = g["cell"]
        assert grids[i].shape == (16, 16, 16)
        assert rel(C64[i], g["C"]) <= 1e-10, (k, rel(C64[i], g["C"]))
        assert rel(Cm[i], g["C"]) <= 1e-6, (k, rel(Cm[i], g["C"]))
