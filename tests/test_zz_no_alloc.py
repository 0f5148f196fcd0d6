"""femh_homogenize_device_noalloc (no device allocation, fixed scratch region)
against femh_homogenize_device with the same options (Jacobi-PCG, FP64):
two calls in a row, the second on a smaller grid, each checked against the
existing entry point's C_H and iteration counts; grids beyond the region's
limit are rejected with BAD_ARG and a message.  Runs in the same process as
the other GPU tests and alone."""
import numpy as np
import pytest

from fem_amd import FemhError, Homogenizer
from fem_amd import workloads as W
from fem_amd.femh import BAD_ARG, NOALLOC_MAX_VOXELS, homogenize_device_noalloc

pytestmark = pytest.mark.gpu
NU = 0.3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a).ravel() - np.asarray(b).ravel()) / np.linalg.norm(b))


def test_noalloc_matches_device_entry_point_on_two_calls():
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    cases = [(20, 24, 28, 3), (12, 12, 12, 4)]  # second call smaller
    kw = dict(nu=NU, penal=3.0, precond="jacobi", tol=1e-10, max_iter=20000)
    for nx, ny, nz, seed in cases:
        shape = (nx, ny, nz)
        assert nx * ny * nz <= NOALLOC_MAX_VOXELS
        rho = W.simp_density(shape, seed, 3.0)
        d_rho = torch.from_numpy(rho).to(dev)
        d_C = torch.zeros(36, dtype=torch.float64, device=dev)
        d_Cref = torch.zeros(36, dtype=torch.float64, device=dev)
        h = Homogenizer(shape, **kw)
        st_ref = h.homogenize_device(d_rho.data_ptr(), d_Cref.data_ptr())
        h.close()
        torch.cuda.synchronize()
        st = homogenize_device_noalloc(shape, d_rho.data_ptr(), d_C.data_ptr(), **kw)
        torch.cuda.synchronize()
        C, Cref = d_C.cpu().numpy(), d_Cref.cpu().numpy()
        print(f"{shape}: rel diff {rel(C, Cref):.2e}, iters {st['iters']} vs {st_ref['iters']}")
        assert rel(C, Cref) <= 1e-12
        assert st["iters"] == st_ref["iters"]
        assert np.isfinite(C).all() and C[0] > 0


def test_noalloc_rejects_oversized_grid_and_other_solvers():
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    d_rho = torch.full((33 * 32 * 32,), 0.5, dtype=torch.float64, device=dev)
    d_C = torch.zeros(36, dtype=torch.float64, device=dev)
    with pytest.raises(FemhError) as e:
        homogenize_device_noalloc((33, 32, 32), d_rho.data_ptr(), d_C.data_ptr(), nu=NU)
    assert e.value.status == BAD_ARG and "32768" in str(e.value)
    with pytest.raises(FemhError) as e:
        homogenize_device_noalloc((8, 8, 8), d_rho.data_ptr(), d_C.data_ptr(), nu=NU, precond="mg")
    assert e.value.status == BAD_ARG
    assert float(d_C.abs().max()) == 0.0
