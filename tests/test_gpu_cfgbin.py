"""GPU test of the tensor-homogenization stage as a batch job over the
sample generator's files (fem_amd/cfgbin.py): tensors and parameters written
in the caller's layout equal one-cell solves of the same cells."""
import numpy as np
import pytest

from fem_amd import Homogenizer, cfgbin

pytestmark = pytest.mark.gpu


def test_run_tensor_homogenization_matches_one_cell_solves(tmp_path):
    level, rep, sub, nsamp = 4, 1, 2, 6
    E1, nu1 = 290.9, 0.45
    rng = np.random.default_rng(17)
    assign = [rng.integers(0, 2, level ** 3).astype(np.int32) for _ in range(nsamp)]
    assign[0][:] = 1  # a homogeneous cell
    root = tmp_path / ("level%d_" % level)
    cfgbin.write_vectors(str(root) + "matAssignments.bin", assign, np.int32)
    cfgbin.write_vectors(str(root) + "baseMat.bin", [np.int32([0]), np.int32([1])], np.int32)
    cfgbin.write_vector(str(root) + "params.bin", np.zeros(nsamp * 4), np.float32)
    tens, par = cfgbin.run_tensor_homogenization(str(tmp_path), level, E1, nu1, [0.0, 1.0], rep=rep,
                                                 subdiv=sub, cubic=True, batch=4, tol=1e-12)
    t_file = cfgbin.read_vector(str(root) + "elasticityTensors.bin", np.float32)
    p_file = cfgbin.read_vector(str(root) + "params.bin", np.float32)
    back = cfgbin.read_vectors(str(root) + "matAssignments.bin", np.int32)
    assert t_file.size == 21 * nsamp and p_file.size == 4 * nsamp
    assert all(np.array_equal(a, b) for a, b in zip(back, assign))
    np.testing.assert_array_equal(t_file.reshape(nsamp, 21), tens)
    n = level * rep * sub
    h = Homogenizer((n, n, n), nu=nu1, E0=E1, interp="modulus", tol=1e-12)
    tril = np.tril_indices(6)
    for s in range(nsamp):
        E = E1 * np.where(cfgbin.replicate_assignment(assign[s], level, rep, sub) == 1, 1.0, 1e-4)
        C, _ = h.homogenize_modulus(E)
        err = np.abs(tens[s] - C[tril]).max() / abs(C[0, 0])
        assert err <= 1e-6, (s, err)
        assert abs(par[s, 0] - assign[s].mean()) <= 1e-6
        Ec, nuc, Gc = cfgbin.cubic_parameters(C)
        np.testing.assert_allclose(par[s, 1:], [Ec, nuc, Gc], rtol=1e-5)
    # the homogeneous cell: the base material's own (E, nu)
    assert abs(par[0, 1] - E1) <= 1e-5 * E1 and abs(par[0, 2] - nu1) <= 1e-5
    h.close()
