"""ctypes wrapper of the CPU FP64 oracle (oracle/femh_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg as the checker.  The product package never imports
this module.

Every function takes and returns numpy float64 arrays in the conventions of
femh_oracle.h: voxel index (i*ny+j)*nz+k, fields component-major [3, V],
load cases in Voigt order (xx, yy, zz, yz, xz, xy) with engineering shear.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _dp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class SolverOpts(ctypes.Structure):
    _fields_ = [("precond", ctypes.c_int), ("levels", ctypes.c_int),
                ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
                ("omega", ctypes.c_double), ("smoother", ctypes.c_int),
                ("cheby_ratio", ctypes.c_double), ("coarse_iters", ctypes.c_int),
                ("tol", ctypes.c_double), ("max_iter", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.or_wall.restype = ctypes.c_double
        L.or_mg_create.restype = ctypes.c_void_p
        L.or_mg_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        for name in ("or_mg_destroy",):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("or_mg_apply", "or_mg_restrict", "or_mg_prolong_add"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.or_mg_level_dims.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _LIB = L
    return _LIB


def _E(E):
    E = np.ascontiguousarray(E, dtype=np.float64)
    assert E.ndim == 3
    return E


def ke_hex(E=1.0, nu=0.3, h=1.0, float_gauss=False):
    K = np.zeros(576)
    lib().or_ke_hex(ctypes.c_double(E), ctypes.c_double(nu), ctypes.c_double(h),
                    ctypes.c_int(int(float_gauss)), _dp(K))
    return K.reshape(24, 24)


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def apply(E, nu, u):
    E = _E(E)
    Ke = ke_hex(1.0, nu).ravel().copy()
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(3, -1)
    y = np.zeros_like(u)
    lib().or_apply(*[ctypes.c_int(n) for n in E.shape], _dp(E), _dp(Ke), _dp(u), _dp(y))
    return y


def diag(E, nu):
    E = _E(E)
    Ke = ke_hex(1.0, nu).ravel().copy()
    d = np.zeros((3, E.size))
    lib().or_diag(*[ctypes.c_int(n) for n in E.shape], _dp(E), _dp(Ke), _dp(d))
    return d


def rhs(E, nu, lc):
    E = _E(E)
    Ke = ke_hex(1.0, nu).ravel().copy()
    b = np.zeros((3, E.size))
    lib().or_rhs(*[ctypes.c_int(n) for n in E.shape], _dp(E), _dp(Ke), ctypes.c_int(lc), _dp(b))
    return b


def opts(precond="mg", levels=16, nu1=2, nu2=2, omega=1.0, smoother="cheby",
         cheby_ratio=8.0, coarse_iters=64, tol=1e-10, max_iter=5000):
    return SolverOpts(1 if precond == "mg" else 0, levels, nu1, nu2, omega,
                      1 if smoother == "cheby" else 0, cheby_ratio, coarse_iters, tol, max_iter)


def homogenize(E, nu, **kw):
    """Returns dict(C=6x6, u=[6,3,V] fluctuation fields, iters, relres, status)."""
    E = _E(E)
    V = E.size
    o = opts(**kw)
    C = np.zeros(36)
    u = np.zeros((6, 3, V))
    it = np.zeros(6, dtype=np.int32)
    rr = np.zeros(6)
    st = lib().or_homogenize(*[ctypes.c_int(n) for n in E.shape], _dp(E), ctypes.c_double(nu),
                             ctypes.byref(o), _dp(C), _dp(u), _dp(it), _dp(rr))
    return dict(C=C.reshape(6, 6), u=u, iters=it, relres=rr, status=st)


def homogenize_timed(E, nu, case_mask=63, **kw):
    """homogenize() for the load cases in case_mask, with the wall seconds of
    the set-up, of those solves and of C_H (dict key "seconds")."""
    E = _E(E)
    V = E.size
    o = opts(**kw)
    C = np.zeros(36)
    u = np.zeros((6, 3, V))
    it = np.zeros(6, dtype=np.int32)
    rr = np.zeros(6)
    sec = np.zeros(3)
    st = lib().or_homogenize_cases(*[ctypes.c_int(n) for n in E.shape], _dp(E), ctypes.c_double(nu),
                                   ctypes.byref(o), ctypes.c_int(case_mask), _dp(C), _dp(u), _dp(it),
                                   _dp(rr), _dp(sec))
    return dict(C=C.reshape(6, 6), u=u, iters=it, relres=rr, status=st, seconds=sec)


def ch_energy(E, nu, u):
    E = _E(E)
    u = np.ascontiguousarray(u, dtype=np.float64)
    C = np.zeros(36)
    lib().or_ch_energy(*[ctypes.c_int(n) for n in E.shape], _dp(E), ctypes.c_double(nu), _dp(u), _dp(C))
    return C.reshape(6, 6)


def thermal_stress(E, nu, alpha, u):
    """beta[6]: effective thermal stress per unit temperature rise from the six
    elastic fields (Gauss-point strains, Lame stress)."""
    E = _E(E)
    a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(E.shape)
    u = np.ascontiguousarray(u, dtype=np.float64)
    beta = np.zeros(6)
    lib().or_thermal_stress(*[ctypes.c_int(n) for n in E.shape], _dp(E), ctypes.c_double(nu), _dp(a),
                            _dp(u), _dp(beta))
    return beta


def energy_terms(shape, nu, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    V = int(np.prod(shape))
    W = np.zeros((V, 21))
    lib().or_energy_terms(*[ctypes.c_int(n) for n in shape], ctypes.c_double(nu), _dp(u), _dp(W))
    return W


def grad_E(shape, nu, u, G):
    u = np.ascontiguousarray(u, dtype=np.float64)
    G = np.ascontiguousarray(G, dtype=np.float64).reshape(36)
    V = int(np.prod(shape))
    g = np.zeros(V)
    lib().or_grad_E(*[ctypes.c_int(n) for n in shape], ctypes.c_double(nu), _dp(u), _dp(G), _dp(g))
    return g.reshape(shape)


class MG:
    """Galerkin multigrid hierarchy of the oracle (test access to levels)."""

    def __init__(self, E, nu, levels=16):
        self.E = _E(E)
        self.h = lib().or_mg_create(*[ctypes.c_int(n) for n in self.E.shape], _dp(self.E),
                                    ctypes.c_double(nu), ctypes.c_int(levels))
        d = np.zeros(3, dtype=np.int32)
        self.L = lib().or_mg_level_dims(self.h, ctypes.c_int(0), _dp(d))
        self.dims = []
        for l in range(self.L):
            lib().or_mg_level_dims(self.h, ctypes.c_int(l), _dp(d))
            self.dims.append(tuple(int(x) for x in d))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mg_destroy(self.h)
            self.h = None

    def apply(self, l, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        y = np.zeros_like(u)
        lib().or_mg_apply(self.h, ctypes.c_int(l), _dp(u), _dp(y))
        return y

    def restrict(self, l, rf):
        rf = np.ascontiguousarray(rf, dtype=np.float64)
        rc = np.zeros(3 * int(np.prod(self.dims[l + 1])))
        lib().or_mg_restrict(self.h, ctypes.c_int(l), _dp(rf), _dp(rc))
        return rc

    def prolong_add(self, l, ec, ef):
        ec = np.ascontiguousarray(ec, dtype=np.float64)
        ef = np.ascontiguousarray(ef, dtype=np.float64).copy()
        lib().or_mg_prolong_add(self.h, ctypes.c_int(l), _dp(ec), _dp(ef))
        return ef


def set_threads(n):
    lib().or_set_threads(ctypes.c_int(n))


def wall():
    return lib().or_wall()


def force_driven(E, nu, regularize=False, faithful=False):
    """The reference's force-driven path (KKT on the (n+1)^3 mesh, face
    tractions, coarse-strain tensor), dense; small grids only.  faithful: the
    reference's float roundings.  Returns dict(C=6x6, u=[6, 3 NV] vertex
    displacements, status)."""
    E = _E(E)
    nx, ny, nz = E.shape
    NV = (nx + 1) * (ny + 1) * (nz + 1)
    C = np.zeros(36)
    u = np.zeros((6, 3 * NV))
    st = lib().or_force_driven(ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz), _dp(E),
                               ctypes.c_double(nu), ctypes.c_int(int(regularize)),
                               ctypes.c_int(int(faithful)), _dp(C), _dp(u))
    return dict(C=C.reshape(6, 6), u=u, status=st)


def kkt_constraints(shape, faithful=False):
    nx, ny, nz = shape
    NV = (nx + 1) * (ny + 1) * (nz + 1)
    n = lib().or_kkt_constraints(nx, ny, nz, int(faithful), None, 0)
    rows = np.zeros((n, 3 * NV))
    got = lib().or_kkt_constraints(nx, ny, nz, int(faithful), _dp(rows), ctypes.c_int(n))
    assert got == n
    return rows


def coarse_strain_tensor(E, nu, u, faithful=False):
    E = _E(E)
    u = np.ascontiguousarray(u, dtype=np.float64)
    C = np.zeros(36)
    lib().or_coarse_strain_tensor(*[ctypes.c_int(n) for n in E.shape], _dp(E), ctypes.c_double(nu),
                                  ctypes.c_int(int(faithful)), _dp(u), _dp(C))
    return C.reshape(6, 6)


def coarse_strain_from_fields(E, nu, u, float_gauss=False):
    E = _E(E)
    u = np.ascontiguousarray(u, dtype=np.float64)
    C = np.zeros(36)
    lib().or_coarse_strain_from_fields(*[ctypes.c_int(n) for n in E.shape], _dp(E), ctypes.c_double(nu),
                                       ctypes.c_int(int(float_gauss)), _dp(u), _dp(C))
    return C.reshape(6, 6)
