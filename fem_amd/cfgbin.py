"""The sample generator's binary files and its tensor-homogenization stage.

Layout (Utilities/include/cfgUtilities.h:803-870, host byte order -- little
endian on this platform): a vector is an int32 count followed by the raw
elements; a vector of vectors is an int32 count followed by each vector in
that form.  The stage's files are level{L}_matAssignments.bin (int vectors,
one per sample, L^3 material indices each), level{L}_baseMat.bin (int
vectors), level{L}_params.bin (float) and level{L}_elasticityTensors.bin
(float, 21 lower-triangle values of C_H per sample, row i, j <= i)
(SamplesGenerator/SamplesGeneratorImpl.cpp:3826-3838, :3854-3923).

run_tensor_homogenization() is SamplesGeneratorImpl::runTensorHomogenization
(:4613-4631) and run_thermal_analysis() is runThermalAnalysis (:4654-4680), each
as a batched GPU job over the engine; the C++ facade has the same jobs
(femh::runTensorHomogenization, femh::runThermalAnalysis).
"""
import os
import struct

import numpy as np

_I32 = np.dtype("<i4")


def write_vector(path, values, dtype):
    a = np.ascontiguousarray(values, dtype=np.dtype(dtype).newbyteorder("<"))
    with open(path, "wb") as f:
        f.write(struct.pack("<i", a.size))
        f.write(a.tobytes())


def write_vectors(path, vectors, dtype):
    dt = np.dtype(dtype).newbyteorder("<")
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(vectors)))
        for v in vectors:
            a = np.ascontiguousarray(v, dtype=dt).ravel()
            f.write(struct.pack("<i", a.size))
            f.write(a.tobytes())


def _read_one(buf, pos, dt):
    if pos + 4 > len(buf):
        raise ValueError("truncated file: missing count")
    (n,) = struct.unpack_from("<i", buf, pos)
    pos += 4
    if n < 0 or pos + n * dt.itemsize > len(buf):
        raise ValueError("truncated file: %d elements announced" % n)
    a = np.frombuffer(buf, dtype=dt, count=n, offset=pos).copy()
    return a, pos + n * dt.itemsize


def read_vector(path, dtype):
    with open(path, "rb") as f:
        buf = f.read()
    a, _ = _read_one(buf, 0, np.dtype(dtype).newbyteorder("<"))
    return a


def read_vectors(path, dtype):
    dt = np.dtype(dtype).newbyteorder("<")
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 4:
        raise ValueError("truncated file: missing count")
    (nvec,) = struct.unpack_from("<i", buf, 0)
    pos, out = 4, []
    for _ in range(nvec):
        a, pos = _read_one(buf, pos, dt)
        out.append(a)
    return out


def replicate_assignment(assignment, level, rep, subdiv):
    """The L^3 assignment replicated rep times per axis with each cell split
    into subdiv^3 voxels (cfgMaterialUtilities.cpp:920-960)."""
    a = np.asarray(assignment, dtype=np.int64).reshape(level, level, level)
    for ax in range(3):
        a = np.repeat(a, subdiv, axis=ax)
    return np.tile(a, (rep, rep, rep))


def cubic_parameters(C):
    """(E, nu, G) of a cubic tensor (NumericalCoarseningUsingMultiGrid.cpp:327-340)."""
    G = C[3, 3]
    a = C[0, 1] / C[0, 0]
    nu = a / (1 + a)
    return C[0, 1] * (1 + nu) * (1 - 2 * nu) / nu, nu, G


def run_tensor_homogenization(directory, level, E_solid, nu_solid, base_densities, rep=1, subdiv=1,
                              cubic=True, density_min=1e-4, batch=16, **solver):
    """Read level{L}_{params,baseMat,matAssignments}.bin from `directory`,
    homogenize every assignment (material 1 solid at E_solid, other materials
    at density_min * E_solid, as NumericalCoarseningUsingMultiGrid.cpp:186-193,
    :239 does), and write the four files back.  Parameters per sample: cubic:
    (density, E, nu, G); otherwise the density.  cubic also selects the
    engine's cubic-structure mode (two load cases for the cells invariant under
    the axis swaps, all six for the others), as the reference's Cubic type
    selects its CUBIC solve (:211).  Returns (tensors[n, 21], params[n, k])."""
    from .femh import Homogenizer
    root = os.path.join(directory, "level%d_" % level)
    read_vector(root + "params.bin", np.float32)  # present, as the reference requires
    base = read_vectors(root + "baseMat.bin", np.int32)
    assign = read_vectors(root + "matAssignments.bin", np.int32)
    n = level * rep * subdiv
    dens = np.asarray(base_densities, dtype=np.float64)
    if cubic:
        solver.setdefault("structure", "cubic")
    solver.setdefault("streams", 3)  # batches in flight together
    h = Homogenizer((n, n, n), nu=nu_solid, E0=E_solid, interp="modulus", batch=batch, **solver)
    tril = np.tril_indices(6)
    tensors, params = [], []
    try:
        step = batch * solver["streams"]
        for c0 in range(0, len(assign), step):
            chunk = assign[c0:c0 + step]
            E = np.stack([E_solid * np.where(replicate_assignment(a, level, rep, subdiv) == 1, 1.0, density_min)
                          for a in chunk])
            C, _ = h.homogenize_many(E, modulus=True)
            for a, Ck in zip(chunk, C):
                tensors.append(Ck[tril])
                p = [dens[np.asarray(a)].mean()]
                if cubic:
                    p.extend(cubic_parameters(Ck))
                params.append(p)
    finally:
        h.close()
    tensors = np.asarray(tensors, dtype=np.float32).reshape(len(assign), 21)
    params = np.asarray(params, dtype=np.float32).reshape(len(assign), -1)
    write_vector(root + "params.bin", params.ravel(), np.float32)
    write_vectors(root + "baseMat.bin", base, np.int32)
    write_vectors(root + "matAssignments.bin", assign, np.int32)
    write_vector(root + "elasticityTensors.bin", tensors.ravel(), np.float32)
    return tensors, params


def run_thermal_analysis(directory, level, E_solid, nu_solid, rep=1, subdiv=1, density_min=1e-4, batch=16,
                         **solver):
    """SamplesGeneratorImpl::runThermalAnalysis (SamplesGeneratorImpl.cpp:4654-4680)
    as a batched GPU job: read level{L}_matAssignments.bin from `directory`,
    compute each cell's effective thermal expansion (material 1 expanding by
    1 per unit temperature rise, other materials at density_min * E_solid not
    expanding, as ThermalAnalysis.cpp:404-420 loads only material 1), and write
    level{L}_thermalExpansionCoefficients.bin with the mean normal expansion
    (a_xx + a_yy + a_zz) / 3 per cell.  Returns (coefficients[n],
    expansions[n, 6])."""
    from .femh import Homogenizer
    root = os.path.join(directory, "level%d_" % level)
    assign = read_vectors(root + "matAssignments.bin", np.int32)
    n = level * rep * subdiv
    h = Homogenizer((n, n, n), nu=nu_solid, E0=E_solid, interp="modulus", batch=batch, **solver)
    out = []
    try:
        for c0 in range(0, len(assign), batch):
            chunk = assign[c0:c0 + batch]
            solid = np.stack([replicate_assignment(a, level, rep, subdiv) == 1 for a in chunk])
            h.homogenize_many(E_solid * np.where(solid, 1.0, density_min), modulus=True)
            _, a = h.thermal_expansion(solid.astype(np.float64), nb=len(chunk))
            out.extend(np.asarray(a).reshape(len(chunk), 6))
    finally:
        h.close()
    alpha = np.asarray(out, dtype=np.float64).reshape(len(assign), 6)
    coef = alpha[:, :3].mean(axis=1).astype(np.float32)
    write_vector(root + "thermalExpansionCoefficients.bin", coef, np.float32)
    return coef, alpha
