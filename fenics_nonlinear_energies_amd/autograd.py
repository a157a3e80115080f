"""torch.autograd boundary for the energy kernels.

The reference differentiates its energies with JAX (``jax.grad`` for the gradient callback,
``reordered_element_base.py:1613-1633``; ``jax.hessian`` for element Hessians, ``:922-923``).
Here the derivatives are the hand-written kernels: ``energy(x, assembler)`` is a torch
scalar whose backward is the gradient kernel (``fne_gradient``), and ``hvp(x, v, assembler)``
applies the assembled CSR Hessian (``fne_hessian`` + ``fne_spmv``).  So a torch optimiser or
any autograd-based caller can drive the same kernels the Newton driver uses.
"""

from __future__ import annotations

import torch


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, assembler) -> torch.Tensor:
        xd = x.detach().to(dtype=torch.float64).contiguous()
        ctx.assembler = assembler
        ctx.save_for_backward(xd)
        return assembler.energy_device(xd, out=torch.empty(1, dtype=torch.float64, device=xd.device))[0]

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (xd,) = ctx.saved_tensors
        g = torch.empty_like(xd)
        ctx.assembler.gradient_device(xd, g)
        return grad_out * g, None


def energy(x: torch.Tensor, assembler) -> torch.Tensor:
    """Total energy J(x) = sum_e E_e - f.x as a differentiable torch scalar (FP64)."""
    return _Energy.apply(x, assembler)


def hvp(x: torch.Tensor, v: torch.Tensor, assembler) -> torch.Tensor:
    """Hessian-vector product H(x) v with the assembled CSR Hessian (assembles at x)."""
    xd = x.detach().to(dtype=torch.float64).contiguous()
    assembler.hessian_device(xd)
    out = torch.empty_like(xd)
    assembler.A.mult_tensor(v.detach().to(dtype=torch.float64).contiguous(), out)
    return out
