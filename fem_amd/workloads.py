"""Seeded synthetic microstructures for the homogenization workloads.

The generator is project-written and portable (SplitMix64 counter hash, then
Box-Muller), so the same seed gives the same voxels on any host.  Arrays are
float64 of shape (nx, ny, nz) in the engine's voxel order (z fastest).

Workloads (sizes and laws as stated for each):
  binary_cube      i.i.d. Bernoulli(p) solid/void voxels
  simp_grf         rho = clamp(0.5 + 0.5 g, 1e-3, 1), g a periodic Gaussian
                   random field (white noise filtered by a Gaussian kernel of
                   standard deviation `corr` voxels, normalised to mean 0 and
                   variance 1); E = E0 * rho**p
  cubic_average    symmetrisation over the cubic group (48 axis permutations
                   and reflections), for the cubic-structure mode
  batch_uneven     cubic grids with edges drawn from {16,24,32,40,48}, fill
                   fraction uniform in [0.2, 0.8], binary voxels
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform(seed, n, stream=0):
    """n uniforms in [0,1) from (seed, stream), counter based."""
    with np.errstate(over="ignore"):
        base = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        ctr = np.arange(n, dtype=np.uint64) + base
        x = _splitmix64(ctr)
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n, stream=0):
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, stream)
    u1 = np.maximum(u[:m], 1e-300)
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def binary_cube(n, seed, p=0.5, stream=0):
    """0/1 solid indicator, shape (n,n,n) or n=(nx,ny,nz)."""
    shape = (n, n, n) if np.isscalar(n) else tuple(n)
    u = uniform(seed, int(np.prod(shape)), stream).reshape(shape)
    return (u < p).astype(np.float64)


def grf(n, seed, corr, stream=0):
    """Periodic Gaussian random field, mean 0, variance 1."""
    shape = (n, n, n) if np.isscalar(n) else tuple(n)
    w = normal(seed, int(np.prod(shape)), stream).reshape(shape)
    f = [np.fft.fftfreq(s) for s in shape]
    k2 = f[0][:, None, None] ** 2 + f[1][None, :, None] ** 2 + f[2][None, None, :] ** 2
    g = np.fft.ifftn(np.fft.fftn(w) * np.exp(-2.0 * np.pi ** 2 * corr ** 2 * k2)).real
    g -= g.mean()
    sd = g.std()
    return g / sd if sd > 0 else g


def simp_density(n, seed, corr=4.0, rho_min=1e-3, stream=0):
    g = grf(n, seed, corr, stream)
    return np.clip(0.5 + 0.5 * g, rho_min, 1.0)


def cubic_average(a):
    """Mean of `a` over the 48 axis permutations and reflections of its last
    three (equal) axes: a cell invariant under the cubic group, the input the
    engine's cubic-structure mode expects."""
    import itertools
    a = np.asarray(a, dtype=np.float64)
    lead = tuple(range(a.ndim - 3))
    acc = np.zeros(a.shape)
    for perm in itertools.permutations(range(3)):
        b = np.transpose(a, lead + tuple(len(lead) + q for q in perm))
        for flips in itertools.product((False, True), repeat=3):
            ax = tuple(len(lead) + q for q, f in enumerate(flips) if f)
            acc += np.flip(b, axis=ax) if ax else b
    return acc / 48.0


def simp_modulus(rho, p=3.0, E0=1.0, Emin=0.0):
    return Emin + (E0 - Emin) * rho ** p


def binary_modulus(solid, E_solid=1.0, E_void=1e-9):
    return np.where(solid > 0.5, E_solid, E_void)


EDGES = (16, 24, 32, 40, 48)


def batch_uneven(count, seed, edges=EDGES, fill=(0.2, 0.8), E_void=1e-9):
    """List of modulus grids (float64) for the uneven batched workload."""
    pick = uniform(seed, count, stream=1)
    frac = uniform(seed, count, stream=2)
    out = []
    for s in range(count):
        n = edges[min(int(pick[s] * len(edges)), len(edges) - 1)]
        f = fill[0] + (fill[1] - fill[0]) * frac[s]
        solid = binary_cube(n, seed, p=f, stream=1000 + s)
        out.append(binary_modulus(solid, 1.0, E_void))
    return out
