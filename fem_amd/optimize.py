"""Batched projected-gradient material optimisation over the C ABI.

Per cell this is the loop of fast_fem/run3d.cpp:238-289 (gradientDescent):
  g = dJ/drho at x; h = max_step / ||g||_inf; line search x' = clamp(x - h g,
  lb, ub), halving h up to 10 times until J(x') < J(x) (run3d.cpp:238-262),
with the objective of fast_fem/FEM3DFun.cpp:198-220,
  J = 0.5 sum_ij wG_ij (C_ij - C0_ij)^2 + 0.5 mw (mean(rho) - m0)^2,
and its gradient dJ/drho_e = sum_ij wG_ij (C_ij - C0_ij) dC_ij/drho_e
+ mw (mean(rho) - m0) / V (fast_fem/FEM3DFun.cpp:222-268).  The density
parameter is per voxel (the field is the identity).

All cells advance together: every forward solve is one batched launch
sequence of up to `batch` cells, warm-started from the cells' previous fields;
cells whose line search already succeeded (or gave up) leave the batch.
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .femh import Homogenizer


def objective(C, C0, wG, rho_mean, m0, mw):
    d = C - C0
    return 0.5 * np.sum(wG * d * d, axis=(-2, -1)) + 0.5 * mw * (rho_mean - m0) ** 2


class BatchedDescent:
    """n cells of one grid; x[n, V] densities in [lb, ub]."""

    def __init__(self, shape, C0, wG=None, m0=0.0, mw=0.0, nu=0.3, batch=8, lb=1e-3, ub=1.0,
                 interp="simp", penal=3.0, **solver):
        self.shape = tuple(shape)
        self.C0 = np.asarray(C0, dtype=np.float64)
        self.n = self.C0.shape[0]
        self.wG = np.ones((self.n, 6, 6)) if wG is None else np.broadcast_to(wG, (self.n, 6, 6))
        self.m0, self.mw, self.lb, self.ub = m0, mw, lb, ub
        self.batch = batch
        self.groups = [list(range(g, min(g + batch, self.n))) for g in range(0, self.n, batch)]
        self.ctx = self._open(nu, interp, penal, solver)
        self.solves = 0
        self.est = []
        self.last_C = None  # C_H at the start point of the last step (its forward solve)
        self._pool = ThreadPoolExecutor(max_workers=min(len(self.groups), 4))

    def _open(self, nu, interp, penal, solver):
        # one context per group of `batch` cells keeps each group's fields
        # resident for the warm start of its next solve
        return [Homogenizer(self.shape, nu=nu, batch=len(g), interp=interp, penal=penal,
                            warm_start=True, **solver) for g in self.groups]

    def _group_solve(self, gi, xg):
        """C_H [nb,6,6] and per-cell stats of group gi at densities xg [nb,V]."""
        return self.ctx[gi].homogenize_many(xg.reshape(len(xg), *self.shape))

    def _group_gradient(self, gi, dJdC):
        """sum_ij dJdC_ij dC_ij/drho [nb,V] at the group's last solve."""
        return self.ctx[gi].gradient(dJdC, nb=len(dJdC)).reshape(len(dJdC), -1)

    def evaluate(self, x, cells=None, grad=False):
        """C (and dJ/drho when grad) of the given cells at densities x[cells]."""
        cells = set(range(self.n)) if cells is None else set(cells)
        C = np.zeros((self.n, 6, 6))
        G = np.zeros_like(x) if grad else None

        def solve(gi):
            # the context solves its whole group (slots keep their cells, so
            # the warm start follows each cell); unselected cells ride along
            # at their current point
            grp = self.groups[gi]
            Cg, st = self._group_solve(gi, x[grp])
            g = None
            if grad:
                rho_mean = x[grp].mean(axis=1)
                dJdC = self.wG[grp] * (Cg - self.C0[grp])
                g = self._group_gradient(gi, dJdC)
                g += (self.mw * (rho_mean - self.m0) / x.shape[1])[:, None]
            return Cg, st, g

        todo = [gi for gi, grp in enumerate(self.groups) if any(c in cells for c in grp)]
        # the groups' contexts run at once (each on its own stream; the
        # library call releases the interpreter lock), filling what one
        # group leaves idle
        if len(todo) > 1:
            out = list(self._pool.map(solve, todo))
        else:
            out = [solve(gi) for gi in todo]
        for gi, (Cg, st, g) in zip(todo, out):
            grp = self.groups[gi]
            self.solves += len(grp)
            self.est.extend(s["ch_err_est"] for s in st)
            C[grp] = Cg
            if grad:
                G[grp] = g
        return C, G

    def J(self, C, x):
        return objective(C, self.C0, self.wG, x.mean(axis=1), self.m0, self.mw)

    def step(self, x, max_step=0.5, max_halvings=10):
        """One projected-gradient step for every cell; returns (x_new, J_new, accepted)."""
        C, G = self.evaluate(x, grad=True)
        self.last_C = C
        J0 = self.J(C, x)
        norm = np.abs(G).max(axis=1)
        h = np.where(norm > 0, max_step / np.maximum(norm, 1e-300), 0.0)
        xn = x.copy()
        accepted = np.zeros(self.n, dtype=bool)
        Jn = J0.copy()
        pending = np.flatnonzero(norm > 0)
        for _ in range(max_halvings):
            if pending.size == 0:
                break
            trial = x.copy()
            trial[pending] = np.clip(x[pending] - h[pending, None] * G[pending], self.lb, self.ub)
            trial[accepted] = xn[accepted]
            Ct, _ = self.evaluate(trial, cells=pending)
            Jt = self.J(Ct, trial)
            ok = pending[Jt[pending] < J0[pending]]
            xn[ok] = trial[ok]
            Jn[ok] = Jt[ok]
            accepted[ok] = True
            h[pending] *= 0.5
            pending = np.setdiff1d(pending, ok)
        return xn, Jn, accepted, J0

    def run(self, x, steps, max_step=0.5, log=None):
        hist = []
        for k in range(steps):
            t = time.perf_counter()
            x, Jn, acc, J0 = self.step(x, max_step)
            hist.append(dict(step=k, J0=J0.copy(), J=Jn.copy(), accepted=int(acc.sum()),
                             seconds=time.perf_counter() - t))
            if log:
                log(hist[-1])
        return x, hist

    def close(self):
        self._pool.shutdown()


class DeviceDescent(BatchedDescent):
    """BatchedDescent with the densities, gradients and trial points in device
    memory.  Every solve reads its cells' densities where they lie
    (femh_homogenize_many_device), the adjoint writes dJ/drho there
    (femh_gradient_device) and the line search forms its trial points there
    (femh_descent_trial, the clamp(x - h g) of fast_fem/run3d.cpp:238-265 with
    the same rounding as the host loop); only C_H, each cell's density sum
    and gradient norm and the per-cell step bookkeeping pass through the
    host.  Same trajectory as BatchedDescent: the trial points are the same
    numbers, and the volume term's mean differs only in summation order.

    x is a float64 device tensor [n, V] (torch provides the device memory);
    step() returns the new point in one of two internal buffers, so the
    point it was given may be the buffer it returned last time."""

    def __init__(self, shape, C0, **kw):
        import torch
        super().__init__(shape, C0, **kw)
        self._torch = torch
        self.V = int(np.prod(self.shape))
        dev = torch.device("cuda")
        self.G = torch.zeros((self.n, self.V), dtype=torch.float64, device=dev)
        self._dC = torch.zeros((self.n, 36), dtype=torch.float64, device=dev)
        self._buf = [torch.zeros((self.n, self.V), dtype=torch.float64, device=dev) for _ in range(2)]
        self.gmax = np.zeros(self.n)
        torch.cuda.synchronize()

    def to_device(self, x):
        """A host array [n, V] as a device tensor (the loop's own buffer)."""
        t = self._torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64).reshape(self.n, self.V))
        self._buf[0].copy_(t)
        self._torch.cuda.synchronize()
        return self._buf[0]

    def _rows(self, t, g0):
        return t.data_ptr() + g0 * t.shape[1] * 8

    def _trial(self, x, h, out):
        """out = clamp(x - h G) / x / unchanged per cell (h > 0 / < 0 / == 0);
        returns the cells' sums of out."""
        return self.ctx[0].descent_trial(self.n, x.data_ptr(), self.G.data_ptr(), h, self.lb, self.ub,
                                         out.data_ptr())

    def evaluate(self, x, cells=None, grad=False, sums=None):
        """C [n,6,6] of the given cells at the device densities x (their whole
        groups are solved); with grad, self.G receives dJ/drho (volume term
        from `sums`, each cell's density sum) and self.gmax its norms."""
        cells = set(range(self.n)) if cells is None else set(cells)
        if not isinstance(x, self._torch.Tensor):
            x = self.to_device(x)
        todo = [gi for gi, grp in enumerate(self.groups) if any(c in cells for c in grp)]

        def solve(gi):
            grp = self.groups[gi]
            g0, nb = grp[0], len(grp)
            st = self.ctx[gi].homogenize_many_device(nb, self._rows(x, g0), self._rows(self._dC, g0))
            if grad:
                Cg = self._dC[g0:g0 + nb].cpu().numpy().reshape(nb, 6, 6)
                shift = None
                if self.mw != 0.0:
                    shift = self.mw * (sums[g0:g0 + nb] / self.V - self.m0) / self.V
                dJdC = self.wG[grp] * (Cg - self.C0[grp])
                self.gmax[g0:g0 + nb] = self.ctx[gi].gradient_device(dJdC, self._rows(self.G, g0), nb=nb,
                                                                     shift=shift)
            return st

        out = list(self._pool.map(solve, todo)) if len(todo) > 1 else [solve(gi) for gi in todo]
        C = np.zeros((self.n, 6, 6))
        Cd = self._dC.cpu().numpy().reshape(self.n, 6, 6)
        for gi, st in zip(todo, out):
            grp = self.groups[gi]
            self.solves += len(grp)
            self.est.extend(s["ch_err_est"] for s in st)
            C[grp] = Cd[grp]
        return C, (self.G if grad else None)

    def Jm(self, C, sums):
        return objective(C, self.C0, self.wG, sums / self.V, self.m0, self.mw)

    def step(self, x, max_step=0.5, max_halvings=10):
        """One projected-gradient step for every cell (x: device [n, V]);
        returns (x_new on the device, J_new, accepted, J0)."""
        n = self.n
        zero = np.zeros(n)
        sx = self._trial(x, zero, x)  # (h = 0: x itself, summed)
        C, _ = self.evaluate(x, grad=True, sums=sx)
        self.last_C = C
        J0 = self.Jm(C, sx)
        norm = self.gmax.copy()
        h = np.where(norm > 0, max_step / np.maximum(norm, 1e-300), 0.0)
        xn = self._buf[1] if x.data_ptr() == self._buf[0].data_ptr() else self._buf[0]
        self._trial(x, -np.ones(n), xn)  # xn = x
        accepted = np.zeros(n, dtype=bool)
        Jn = J0.copy()
        pending = np.flatnonzero(norm > 0)
        for _ in range(max_halvings):
            if pending.size == 0:
                break
            hh = zero.copy()
            hh[pending] = h[pending]
            st = self._trial(x, hh, xn)
            Ct, _ = self.evaluate(xn, cells=pending)
            Jt = self.Jm(Ct, st)
            ok = pending[Jt[pending] < J0[pending]]
            Jn[ok] = Jt[ok]
            accepted[ok] = True
            h[pending] *= 0.5
            pending = np.setdiff1d(pending, ok)
        if pending.size:  # no halving lowered J: those cells stay at x
            hh = zero.copy()
            hh[pending] = -1.0
            self._trial(x, hh, xn)
        return xn, Jn, accepted, J0

    def run(self, x, steps, max_step=0.5, log=None):
        if not isinstance(x, self._torch.Tensor):
            x = self.to_device(x)
        return super().run(x, steps, max_step, log)
        for h in self.ctx:
            h.close()
