"""ctypes binding of the engine's C ABI (include/femh.h).

This is the Python face of the same entry points the C++ host facade uses:
density (or modulus) grid in, 6x6 homogenized tensor C_H out, dJ/drho from the
element-energy adjoint.  Arrays are numpy float64, voxel order (i, j, k) with
k fastest.  The library is the in-tree fem_amd/libfemh.so; there is no CPU
fallback -- a missing or unloadable library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfemh.so")
_LIB = None

OK, BAD_ARG, NOT_CONVERGED, NONFINITE, DEVICE_ERROR, OOM, NO_FIELDS = range(7)
_STATUS = {0: "OK", 1: "BAD_ARG", 2: "NOT_CONVERGED", 3: "NONFINITE", 4: "DEVICE_ERROR",
           5: "OOM", 6: "NO_FIELDS"}
INTERP = {"simp": 0, "fastfem": 1, "modulus": 2}
PRECOND = {"jacobi": 0, "mg": 1}
PRECISION = {"fp64": 0, "mixed": 1}

# Every symbol include/femh.h declares.
EXPORTS = ("femh_abi_version", "femh_default_opts", "femh_create", "femh_destroy",
           "femh_homogenize", "femh_homogenize_modulus", "femh_homogenize_device",
           "femh_homogenize_many", "femh_homogenize_many_modulus", "femh_homogenize_many_device",
           "femh_gradient", "femh_gradient_device", "femh_descent_trial",
           "femh_energy_terms", "femh_get_fields", "femh_thermal_expansion",
           "femh_coarse_strain_tensor",
           "femh_launch_shapes", "femh_set_launch_shapes", "femh_homogenize_batch",
           "femh_prepare", "femh_apply_level", "femh_level_dims", "femh_set_profiling",
           "femh_kernel_stats", "femh_element_stiffness", "femh_last_error",
           "femh_homogenize_device_noalloc")
NOALLOC_MAX_VOXELS = 32768


class FemhError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"femh {_STATUS.get(status, status)}: {msg}")
        self.status = status


class Opts(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("E0", ctypes.c_double), ("Emin", ctypes.c_double), ("nu", ctypes.c_double),
                ("penal", ctypes.c_double), ("interp", ctypes.c_int), ("precond", ctypes.c_int),
                ("levels", ctypes.c_int), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
                ("cheby_ratio", ctypes.c_double), ("coarse_iters", ctypes.c_int),
                ("tol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("precision", ctypes.c_int), ("ch_tol", ctypes.c_double),
                ("check_every", ctypes.c_int), ("device", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("max_outer", ctypes.c_int),
                ("apply_kernel", ctypes.c_int), ("batch", ctypes.c_int),
                ("coarse_nu", ctypes.c_int), ("warm_start", ctypes.c_int),
                ("structure", ctypes.c_int), ("autotune", ctypes.c_int),
                ("streams", ctypes.c_int), ("batch_threads", ctypes.c_int),
                ("graph", ctypes.c_int), ("adapt_inner", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int * 6), ("outer", ctypes.c_int * 6),
                ("relres", ctypes.c_double * 6), ("ch_err_est", ctypes.c_double),
                ("ms_setup", ctypes.c_double), ("ms_solve", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("levels", ctypes.c_int),
                ("fp64_fallback", ctypes.c_int), ("ritz_min", ctypes.c_double)]

    def as_dict(self):
        return dict(iters=list(self.iters), outer=list(self.outer), relres=list(self.relres),
                    ch_err_est=self.ch_err_est, ms_setup=self.ms_setup, ms_solve=self.ms_solve,
                    ms_total=self.ms_total, levels=self.levels, fp64_fallback=self.fp64_fallback,
                    ritz_min=self.ritz_min)


def lib():
    """Load libfemh.so (importing torch first when present so that a single
    HIP runtime serves both)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        try:
            import torch  # noqa: F401  (pins torch's libamdhip64 before ours)
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.femh_create.argtypes = [ctypes.POINTER(Opts), ctypes.POINTER(vp)]
        L.femh_destroy.argtypes = [vp]
        L.femh_destroy.restype = None
        L.femh_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.femh_default_opts.restype = None
        for name in ("femh_homogenize", "femh_homogenize_modulus"):
            getattr(L, name).argtypes = [vp, vp, vp, vp]
        L.femh_homogenize_device.argtypes = [vp, vp, vp, vp, vp]
        for name in ("femh_homogenize_many", "femh_homogenize_many_modulus"):
            getattr(L, name).argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.femh_homogenize_many_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.femh_gradient.argtypes = [vp, vp, vp]
        L.femh_gradient_device.argtypes = [vp, vp, vp, vp, vp, vp]
        L.femh_descent_trial.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_double, ctypes.c_double,
                                         vp, vp, vp]
        L.femh_energy_terms.argtypes = [vp, vp]
        L.femh_get_fields.argtypes = [vp, vp]
        L.femh_thermal_expansion.argtypes = [vp, vp, vp, vp]
        L.femh_coarse_strain_tensor.argtypes = [vp, vp]
        L.femh_launch_shapes.argtypes = [vp, vp, ctypes.c_int]
        L.femh_set_launch_shapes.argtypes = [vp, vp, ctypes.c_int]
        L.femh_homogenize_batch.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.femh_prepare.argtypes = [vp, vp]
        L.femh_apply_level.argtypes = [vp, ctypes.c_int, vp, vp]
        L.femh_level_dims.argtypes = [vp, ctypes.c_int, vp]
        L.femh_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.femh_set_profiling.argtypes = [vp, ctypes.c_int]
        L.femh_kernel_stats.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.femh_element_stiffness.argtypes = [ctypes.c_double, ctypes.c_double, vp]
        L.femh_homogenize_device_noalloc.argtypes = [ctypes.POINTER(Opts), vp, vp, vp, vp,
                                                     ctypes.c_char_p, ctypes.c_size_t]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_opts(shape, nu=0.3, E0=1.0, Emin=0.0, penal=3.0, interp="simp", precond="mg",
              levels=0, nu1=2, nu2=2, cheby_ratio=20.0, coarse_iters=32, tol=1e-10,
              max_iter=5000, precision="fp64", ch_tol=0.0, check_every=1, device=0,
              inner_tol=1e-4, max_outer=12, apply_kernel="tiled", batch=1, warm_start=False,
              coarse_nu=3, structure="general", autotune=False, streams=1, batch_threads=0,
              graph=True, adapt_inner=True):
    o = Opts()
    lib().femh_default_opts(ctypes.byref(o))
    o.nx, o.ny, o.nz = (int(s) for s in shape)
    o.E0, o.Emin, o.nu, o.penal = E0, Emin, nu, penal
    o.interp = INTERP[interp]
    o.precond = PRECOND[precond]
    o.levels, o.nu1, o.nu2 = levels, nu1, nu2
    o.cheby_ratio, o.coarse_iters = cheby_ratio, coarse_iters
    o.tol, o.max_iter = tol, max_iter
    o.precision = PRECISION[precision]
    o.ch_tol, o.check_every, o.device = ch_tol, check_every, device
    o.inner_tol, o.max_outer = inner_tol, max_outer
    o.apply_kernel = {"tiled": 0, "dense": 1}[apply_kernel]
    o.batch = batch
    o.warm_start = int(bool(warm_start))
    o.coarse_nu = coarse_nu
    o.structure = {"general": 0, "cubic": 1}[structure]
    o.autotune = int(bool(autotune))
    o.streams = int(streams)
    o.batch_threads = int(batch_threads)
    o.graph = int(bool(graph))
    o.adapt_inner = int(bool(adapt_inner))
    return o


class Homogenizer:
    """One device context for grids of a fixed shape (mirrors one
    NumericalCoarseningUsingMultigrid instance of the reference)."""

    def __init__(self, shape, **kw):
        self.shape = tuple(int(s) for s in shape)
        self.V = int(np.prod(self.shape))
        self.opts = make_opts(self.shape, **kw)
        h = ctypes.c_void_p()
        st = lib().femh_create(ctypes.byref(self.opts), ctypes.byref(h))
        self._h = h
        if st != OK:
            msg = self._err()
            self.close()
            raise FemhError(st, msg)

    def _err(self):
        buf = ctypes.create_string_buffer(512)
        if self._h:
            lib().femh_last_error(self._h, buf, 512)
        return buf.value.decode(errors="replace")

    def _check(self, st):
        if st != OK:
            raise FemhError(st, self._err())

    def close(self):
        if getattr(self, "_h", None):
            lib().femh_destroy(self._h)
            self._h = None

    __del__ = close

    def _grid(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.V:
            raise ValueError(f"expected {self.V} voxels, got {a.size}")
        return a

    def homogenize(self, rho, check=True):
        rho = self._grid(rho)
        C = np.zeros(36)
        s = Stats()
        st = lib().femh_homogenize(self._h, _ptr(rho), _ptr(C), ctypes.byref(s))
        if check:
            self._check(st)
        return C.reshape(6, 6), s.as_dict()

    def homogenize_modulus(self, E, check=True):
        E = self._grid(E)
        C = np.zeros(36)
        s = Stats()
        st = lib().femh_homogenize_modulus(self._h, _ptr(E), _ptr(C), ctypes.byref(s))
        if check:
            self._check(st)
        return C.reshape(6, 6), s.as_dict()

    def homogenize_device(self, d_rho_ptr, d_C_ptr, stream_ptr=None):
        s = Stats()
        self._check(lib().femh_homogenize_device(self._h, ctypes.c_void_p(d_rho_ptr),
                                                 ctypes.c_void_p(d_C_ptr),
                                                 ctypes.c_void_p(stream_ptr or 0), ctypes.byref(s)))
        return s.as_dict()

    def homogenize_many(self, grids, modulus=False, check=True):
        """n grids of this shape (array [n, *shape]), `batch` per launch
        sequence.  Returns (C[n,6,6], list of stats dicts)."""
        a = np.ascontiguousarray(grids, dtype=np.float64).reshape(-1)
        if a.size % self.V:
            raise ValueError("grid count does not divide the array")
        n = a.size // self.V
        C = np.zeros((n, 36))
        S = (Stats * n)()
        fn = lib().femh_homogenize_many_modulus if modulus else lib().femh_homogenize_many
        st = fn(self._h, n, _ptr(a), _ptr(C), S)
        if check:
            self._check(st)
        return C.reshape(n, 6, 6), [s.as_dict() for s in S]

    def homogenize_many_device(self, n, d_rho_ptr, d_C_ptr, stream_ptr=None):
        S = (Stats * n)()
        self._check(lib().femh_homogenize_many_device(self._h, n, ctypes.c_void_p(d_rho_ptr),
                                                      ctypes.c_void_p(d_C_ptr),
                                                      ctypes.c_void_p(stream_ptr or 0), S))
        return [s.as_dict() for s in S]

    def gradient(self, dJdC, nb=1):
        """dJ/drho of the last batch of nb samples (dJdC [nb,6,6])."""
        G = np.ascontiguousarray(dJdC, dtype=np.float64).reshape(36 * nb)
        out = np.zeros(self.V * nb)
        self._check(lib().femh_gradient(self._h, _ptr(G), _ptr(out)))
        return out.reshape(self.shape) if nb == 1 else out.reshape((nb,) + self.shape)

    def gradient_device(self, dJdC, d_out_ptr, nb=1, shift=None, stream_ptr=None):
        """dJ/drho of the last batch of nb samples into device memory at
        d_out_ptr (nb*V doubles), plus shift[s] per sample when given.
        Returns max |dJ/drho| per sample."""
        G = np.ascontiguousarray(dJdC, dtype=np.float64).reshape(36 * nb)
        sh = None if shift is None else np.ascontiguousarray(shift, dtype=np.float64).reshape(nb)
        gmax = np.zeros(nb)
        self._check(lib().femh_gradient_device(self._h, _ptr(G), None if sh is None else _ptr(sh),
                                               ctypes.c_void_p(d_out_ptr), _ptr(gmax),
                                               ctypes.c_void_p(stream_ptr or 0)))
        return gmax

    def descent_trial(self, n, d_x_ptr, d_g_ptr, h, lb, ub, d_out_ptr, stream_ptr=None):
        """Projected-gradient trial points of n cells of this grid in device
        memory: per cell, h > 0: out = clamp(x - h g, lb, ub); h < 0: out = x;
        h == 0: out unchanged.  Returns each cell's sum of out."""
        hh = np.ascontiguousarray(h, dtype=np.float64).reshape(n)
        sums = np.zeros(n)
        self._check(lib().femh_descent_trial(self._h, n, ctypes.c_void_p(d_x_ptr), ctypes.c_void_p(d_g_ptr),
                                             _ptr(hh), float(lb), float(ub), ctypes.c_void_p(d_out_ptr),
                                             _ptr(sums), ctypes.c_void_p(stream_ptr or 0)))
        return sums

    def energy_terms(self, nb=1):
        W = np.zeros((nb * self.V, 21))
        self._check(lib().femh_energy_terms(self._h, _ptr(W)))
        return W if nb == 1 else W.reshape(nb, self.V, 21)

    def fields(self, nb=1):
        u = np.zeros((nb * 6, 3, self.V))
        self._check(lib().femh_get_fields(self._h, _ptr(u)))
        return u if nb == 1 else u.reshape(nb, 6, 3, self.V)

    def thermal_expansion(self, alpha, nb=1):
        """Effective thermal stress beta and expansion C_H^-1 beta (Voigt,
        engineering shear) of the last batch of nb cells, for the per-voxel
        expansion coefficients alpha ([nb,] nx, ny, nz).  Returns (beta,
        alpha_eff), each [6] (nb = 1) or [nb, 6]."""
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(nb * self.V)
        beta = np.zeros(6 * nb)
        aeff = np.zeros(6 * nb)
        self._check(lib().femh_thermal_expansion(self._h, _ptr(a), _ptr(beta), _ptr(aeff)))
        if nb == 1:
            return beta, aeff
        return beta.reshape(nb, 6), aeff.reshape(nb, 6)

    def coarse_strain_tensor(self, nb=1):
        """The force-driven (coarse-strain) form of C_H of the last batch of
        nb cells: [6, 6] (nb = 1) or [nb, 6, 6]."""
        C = np.zeros(36 * nb)
        self._check(lib().femh_coarse_strain_tensor(self._h, _ptr(C)))
        return C.reshape(6, 6) if nb == 1 else C.reshape(nb, 6, 6)

    def launch_shapes(self):
        """Launch shapes in use: [fine-operator planes per workgroup, transfer
        planes per thread of each level pair]; 0 = the built-in choice."""
        buf = (ctypes.c_int * 16)()
        n = lib().femh_launch_shapes(self._h, buf, 16)
        if n < 0:
            self._check(-n)
        return [buf[i] for i in range(min(n, 16))]

    def set_launch_shapes(self, shapes):
        """Pin the launch shapes (femh_launch_shapes encoding; 0 = built-in)."""
        a = np.ascontiguousarray(shapes, dtype=np.int32)
        self._check(lib().femh_set_launch_shapes(self._h, _ptr(a), int(a.size)))

    def set_profiling(self, on=True):
        self._check(lib().femh_set_profiling(self._h, int(bool(on))))

    def kernel_stats(self, which):
        """which: 0 = FP64 fine apply, 1 = FP32 fine apply."""
        n = ctypes.c_longlong()
        ms, by, fl = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().femh_kernel_stats(self._h, which, ctypes.byref(n), ctypes.byref(ms),
                                            ctypes.byref(by), ctypes.byref(fl)))
        return dict(launches=n.value, ms=ms.value, bytes=by.value, flops=fl.value)

    def prepare(self, E):
        self._check(lib().femh_prepare(self._h, _ptr(self._grid(E))))

    def level_dims(self, level):
        d = np.zeros(3, dtype=np.int32)
        st = lib().femh_level_dims(self._h, level, _ptr(d))
        return None if st != OK else tuple(int(v) for v in d)

    def apply_level(self, level, u):
        d = self.level_dims(level)
        if d is None:
            raise FemhError(BAD_ARG, f"no level {level}")
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if u.size != 3 * int(np.prod(d)):
            raise ValueError("field size does not match the level")
        y = np.zeros_like(u)
        self._check(lib().femh_apply_level(self._h, level, _ptr(u), _ptr(y)))
        return y


def homogenize_device_noalloc(shape, d_rho_ptr, d_C_ptr, stream_ptr=None, **kw):
    """One cell through femh_homogenize_device_noalloc (Jacobi-PCG, FP64; device
    pointers).  kw are make_opts() options; precond defaults to "jacobi".
    Returns the stats dict; raises FemhError on failure."""
    kw.setdefault("precond", "jacobi")
    o = make_opts(shape, **kw)
    s = Stats()
    buf = ctypes.create_string_buffer(512)
    st = lib().femh_homogenize_device_noalloc(ctypes.byref(o), ctypes.c_void_p(d_rho_ptr),
                                              ctypes.c_void_p(d_C_ptr), ctypes.c_void_p(stream_ptr or 0),
                                              ctypes.byref(s), buf, 512)
    if st != OK:
        raise FemhError(st, buf.value.decode(errors="replace"))
    return s.as_dict()


def element_stiffness(E=1.0, nu=0.3):
    """24x24 unit-voxel stiffness from the engine's closed-form tables (host only)."""
    K = np.zeros(576)
    st = lib().femh_element_stiffness(E, nu, _ptr(K))
    if st != OK:
        raise FemhError(st, "bad material")
    return K.reshape(24, 24)


class BatchHomogenizer:
    """Uneven-size batches over modulus grids (femh_homogenize_batch) through
    one context whose per-size sub-contexts stay resident between calls, so
    that repeated batches pay their set-up once."""

    def __init__(self, **kw):
        self._h = Homogenizer((8, 8, 8), interp="modulus", **kw)

    def run(self, grids):
        """Returns (C[n,6,6], stats)."""
        grids = [np.ascontiguousarray(g, dtype=np.float64) for g in grids]
        n = len(grids)
        if n == 0:
            return np.zeros((0, 6, 6)), {}
        dims = np.array([g.shape for g in grids], dtype=np.int32).reshape(-1)
        ptrs = (ctypes.c_void_p * n)(*[g.ctypes.data for g in grids])
        C = np.zeros((n, 36))
        s = Stats()
        self._h._check(lib().femh_homogenize_batch(self._h._h, n, _ptr(dims), ptrs, _ptr(C), ctypes.byref(s)))
        return C.reshape(n, 6, 6), s.as_dict()

    def close(self):
        self._h.close()


def homogenize_batch(grids, **kw):
    """Batch over modulus grids of uneven shapes.  Returns (C[n,6,6], stats)."""
    b = BatchHomogenizer(**kw)
    try:
        return b.run(grids)
    finally:
        b.close()
