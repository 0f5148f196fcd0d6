"""torch.autograd.Function over the engine's C ABI: rho -> C_H with the
element-energy adjoint as backward (fast_fem/FEM3DFun.cpp:222-268 computes
the same dG_ij/dx_e = coef'(x_e) u_i^T K0 u_j for its objective)."""
import numpy as np
import torch

from .femh import Homogenizer


class HomogenizeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, hom):
        r = rho.detach().to("cpu", torch.float64).contiguous().numpy().copy()
        C, _ = hom.homogenize(r)
        hom._last_rho = r
        ctx.hom = hom
        ctx.rho = r
        return torch.from_numpy(np.array(C)).to(dtype=rho.dtype, device=rho.device)

    @staticmethod
    def backward(ctx, gC):
        hom = ctx.hom
        if getattr(hom, "_last_rho", None) is None or not np.array_equal(hom._last_rho, ctx.rho):
            hom.homogenize(ctx.rho)  # the context holds the fields of its last solve only
            hom._last_rho = ctx.rho
        g = hom.gradient(gC.detach().to("cpu", torch.float64).numpy())
        return torch.from_numpy(g).to(dtype=gC.dtype, device=gC.device), None


def homogenize(rho: torch.Tensor, hom: Homogenizer) -> torch.Tensor:
    """Differentiable C_H(rho) for a density grid shaped like the context."""
    return HomogenizeFn.apply(rho, hom)
