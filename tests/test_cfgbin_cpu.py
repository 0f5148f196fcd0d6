"""The sample generator's binary file layout (fem_amd/cfgbin.py), on the CPU:
exact bytes of the int32-count + raw-elements form, round trips, truncated
files, and the material-assignment replication."""
import struct

import numpy as np
import pytest

from fem_amd import cfgbin


def test_vector_bytes_and_round_trip(tmp_path):
    p = tmp_path / "v.bin"
    cfgbin.write_vector(p, [1.5, -2.0, 3.25], np.float32)
    assert p.read_bytes() == struct.pack("<i", 3) + struct.pack("<3f", 1.5, -2.0, 3.25)
    np.testing.assert_array_equal(cfgbin.read_vector(p, np.float32), np.float32([1.5, -2.0, 3.25]))
    cfgbin.write_vector(p, [], np.int32)
    assert p.read_bytes() == struct.pack("<i", 0)
    assert cfgbin.read_vector(p, np.int32).size == 0


def test_vector_of_vectors_bytes_and_round_trip(tmp_path):
    p = tmp_path / "vv.bin"
    vv = [np.int32([0, 1, 1]), np.int32([]), np.int32([7])]
    cfgbin.write_vectors(p, vv, np.int32)
    expect = struct.pack("<i", 3) + struct.pack("<i3i", 3, 0, 1, 1) + struct.pack("<i", 0) + struct.pack("<ii", 1, 7)
    assert p.read_bytes() == expect
    back = cfgbin.read_vectors(p, np.int32)
    assert len(back) == 3 and all(np.array_equal(a, b) for a, b in zip(back, vv))


def test_truncated_files_are_rejected(tmp_path):
    p = tmp_path / "t.bin"
    p.write_bytes(struct.pack("<i", 4) + struct.pack("<3f", 1, 2, 3))
    with pytest.raises(ValueError):
        cfgbin.read_vector(p, np.float32)
    p.write_bytes(struct.pack("<i", 2) + struct.pack("<ii", 1, 5))
    with pytest.raises(ValueError):
        cfgbin.read_vectors(p, np.int32)
    p.write_bytes(b"\x01\x00")
    with pytest.raises(ValueError):
        cfgbin.read_vectors(p, np.int32)


def test_replicate_assignment_matches_index_formula():
    rng = np.random.default_rng(3)
    L, rep, sub = 3, 2, 2
    a = rng.integers(0, 2, L ** 3)
    out = cfgbin.replicate_assignment(a, L, rep, sub)
    n = L * rep * sub
    assert out.shape == (n, n, n)
    # out[X, Y, Z] = in[(X / sub) % L, (Y / sub) % L, (Z / sub) % L]
    for X in range(n):
        for Y in range(n):
            for Z in range(n):
                assert out[X, Y, Z] == a[(((X // sub) % L) * L + (Y // sub) % L) * L + (Z // sub) % L]


def test_cubic_parameters_of_isotropic_tensor():
    E, nu = 2.5, 0.3
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    Ec, nuc, Gc = cfgbin.cubic_parameters(C)
    assert abs(Ec - E) < 1e-12 and abs(nuc - nu) < 1e-12 and abs(Gc - mu) < 1e-12
