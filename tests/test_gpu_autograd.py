"""torch.autograd boundary: backward of the energy is the gradient kernel; the Hessian-vector
product is the assembled CSR Hessian.  Checked against central finite differences of the
energy kernel (h = 1e-6 along seeded random directions, rel. 1e-6) and, for the gradient,
bitwise against fne_gradient."""

import numpy as np
import pytest
import torch

from fenics_nonlinear_energies_amd.autograd import energy, hvp
from fenics_nonlinear_energies_amd.problems.ginzburg_landau import GLReorderedElementAssembler, build_gl_problem
from fenics_nonlinear_energies_amd.problems.hyperelasticity import HEReorderedElementAssembler, build_beam_problem

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind == "gl":
        p, adj = build_gl_problem(32)
        asm = GLReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
        u = np.random.default_rng(0).uniform(-1, 1, asm.layout.n_free)
    else:
        p, adj, u0 = build_beam_problem(12, 2, 2)
        asm = HEReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
        u = (u0 + 1e-4 * np.random.default_rng(0).standard_normal(u0.size))[asm.layout.perm]
    return asm, torch.as_tensor(u, device="cuda")


@pytest.mark.parametrize("kind", ["gl", "he"])
def test_energy_backward_and_hvp(require_gpu, kind):
    asm, x = _case(kind)
    xr = x.clone().requires_grad_(True)
    E = energy(xr, asm)
    (g,) = torch.autograd.grad(E, xr)
    gk = torch.empty_like(x)
    asm.gradient_device(x, gk)
    assert torch.equal(g, gk)
    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(3):
        d = torch.as_tensor(rng.standard_normal(x.numel()), device="cuda")
        d = d / d.norm() * x.norm().clamp_min(1.0) * 1e-2
        fd = (float(energy(x + h * d, asm)) - float(energy(x - h * d, asm))) / (2 * h)
        an = float(g @ d)
        assert abs(fd - an) <= 1e-6 * max(abs(an), float(g.norm() * d.norm()))
        Hd = hvp(x, d, asm)
        gp, gm = torch.empty_like(x), torch.empty_like(x)
        asm.gradient_device(x + h * d, gp)
        asm.gradient_device(x - h * d, gm)
        fdh = (gp - gm) / (2 * h)
        assert float((fdh - Hd).norm()) <= 1e-5 * float(Hd.norm())
    asm.cleanup()
