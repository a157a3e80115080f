"""Parity at the full sizes of the benchmark workloads.

* Ginzburg-Landau 1024^2 (2,097,152 P1 elements, 1,046,529 free DOFs, the headline workload with
  its sine-plus-noise state, seed 0): the fused assembly against the C restatement of the oracle
  (``oracle/fne_oracle.c``, FP64), FP64 and FP32 Hessian values.  Tolerances: energy 1e-11
  relative (2M terms summed in different orders), gradient 1e-12 of its largest entry, FP64
  Hessian 1e-12 and FP32 Hessian 2e-6 of the largest entry (a row entry sums at most six FP32
  element values in FP32).
* Neo-Hookean beam 320x16x16 (491,520 tets, 276,573 free DOFs): size-independent properties of
  the fused FP64 assembly at a deformed state -- the CSR Hessian is symmetric to 1e-13 of its
  largest entry (the element matrices G^T A G are symmetric up to rounding), the
  gradient is the central difference of the energy along a random direction and H d the central
  difference of the gradient (rel. 1e-5 each at h = 1e-6 relative: the O(h^2) remainder plus
  rounding; a dropped term or a wrong sign is off by O(1)), and a repeated assembly gives the
  same bits.
* Mohr-Coulomb P4 slope at h = 1/7 (98,000 elements, 1,568,280 free DOFs) in a plastic state:
  the assembled FP64 Hessian's product equals the matrix-free product fne_hvp (rel. 1e-10 of the
  product's largest entry; the reference pins the same equality at 1e-8,
  tests/test_slope_stability_petsc.py:475-515), the Hessian is symmetric to 1e-12 of its largest
  entry, and the gradient is the central difference of the energy (rel. 1e-6).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_gl_headline_size_matches_the_c_oracle(require_gpu, prec):
    from fenics_nonlinear_energies_amd.problems.ginzburg_landau import (
        GLReorderedElementAssembler,
        build_gl_problem,
        sine_noise_init,
    )
    from oracle import c_assembly

    if not c_assembly._LIB.exists():
        pytest.skip("oracle/_lib/libfne_oracle.so not built")
    params, adj = build_gl_problem(1024)
    u = sine_noise_init(params, 0)
    asm = GLReorderedElementAssembler(params, None, adj, hessian_precision=prec)
    lay = asm.layout
    x = asm.create_vec(u[lay.perm])
    g = x.duplicate()
    e_gpu = float(asm.energy_grad_hessian_device(x.t, g.t).item())
    H = asm.A.to_scipy()[lay.iperm][:, lay.iperm].tocsr()
    gg = g.array[lay.iperm]
    asm.cleanup()
    E, gref, Href = c_assembly.P1Assembly("gl", params)(u)
    print(f"GL 1024^2 {prec}: E {e_gpu!r} oracle {E!r}", flush=True)
    assert _rel(e_gpu, E) <= 1e-11
    assert np.abs(gg - gref).max() <= 1e-12 * np.abs(gref).max()
    H.sort_indices()
    assert H.nnz == Href.nnz == 7317521
    assert np.array_equal(H.indptr, Href.indptr) and np.array_equal(H.indices, Href.indices)
    tol = 2e-6 if prec == "fp32" else 1e-12
    assert np.abs(H.data - Href.data).max() <= tol * np.abs(Href.data).max()


def test_he_c4_size_symmetry_finite_differences_determinism(require_gpu):
    from fenics_nonlinear_energies_amd.problems.hyperelasticity import (
        HEReorderedElementAssembler,
        build_beam_problem,
        rotate_right_face_from_reference,
    )

    p, adj, u0 = build_beam_problem(320, 16, 16)
    asm = HEReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
    asm.update_dirichlet(rotate_right_face_from_reference(p["u_0_ref"], p["nodes2coord"], 0.02, p["right_nodes"]))
    rng = np.random.default_rng(5)
    u = u0 + 2e-5 * rng.standard_normal(u0.size)
    x = asm.create_vec(u[asm.layout.perm])
    g = x.duplicate()
    asm.energy_grad_hessian_device(x.t, g.t)
    vals1 = asm.A.data.clone()
    asm.energy_grad_hessian_device(x.t, g.t)
    assert torch.equal(vals1, asm.A.data)
    H = asm.A.to_scipy()
    print(f"HE 320x16x16: max |H - H^T| / max |H| = {abs(H - H.T).max() / abs(H).max():.3e}", flush=True)
    assert abs(H - H.T).max() <= 1e-13 * abs(H).max()
    d = torch.as_tensor(rng.standard_normal(x.t.numel()), device=x.t.device)
    h = 1e-6 * float(torch.linalg.vector_norm(x.t)) / float(torch.linalg.vector_norm(d))
    ep = float(asm.energy_device(x.t + h * d).item())
    em = float(asm.energy_device(x.t - h * d).item())
    gd = float(torch.dot(g.t, d))
    print(f"HE 320x16x16: (E+ - E-)/2h {(ep - em) / (2 * h)!r} g.d {gd!r}", flush=True)
    assert _rel((ep - em) / (2 * h), gd) <= 1e-5
    gp, gm = torch.empty_like(x.t), torch.empty_like(x.t)
    asm.gradient_device(x.t + h * d, gp)
    asm.gradient_device(x.t - h * d, gm)
    fd = (gp - gm) / (2 * h)
    hd = torch.empty_like(x.t)
    asm.A.mult_tensor(d, hd)
    rel_hd = float(torch.linalg.vector_norm(fd - hd)) / float(torch.linalg.vector_norm(hd))
    print(f"HE 320x16x16: |FD of g - H d| / |H d| = {rel_hd:.3e}", flush=True)
    assert rel_hd <= 1e-5
    asm.cleanup()


def test_p4_slope_c5_size_matrix_free_product_symmetry_gradient(require_gpu):
    from fenics_nonlinear_energies_amd.problems import slope_stability as S

    geo = dict(x1=20.0, x2=20.0, x3=20.0, y1=10.0, y2=10.0)
    mat = dict(E=20000.0, nu=0.49, c0=6.0, phi_deg=45.0, psi_deg=0.0, gamma=20.0)
    p, _, _ = S.build_slope_problem(1.0 / 7.0, degree=4, geometry=geo, material=mat, with_adjacency=False)
    asm = S.SlopeStabilityReorderedElementAssembler(p, None, None, hessian_precision="fp64")
    X = p["nodes"]
    v = np.zeros(2 * X.shape[0])
    v[0::2] = 1e-3 * X[:, 1]  # shear growing with height: plastic near the toe, elastic elsewhere
    v[1::2] = -5e-4 * X[:, 1]
    x = asm.create_vec(v[p["freedofs"]][asm.layout.perm])
    g = x.duplicate()
    asm.energy_grad_hessian_device(x.t, g.t)
    rng = np.random.default_rng(9)
    d = torch.as_tensor(rng.standard_normal(x.t.numel()), device=x.t.device)
    y_csr, y_mf = torch.empty_like(x.t), torch.empty_like(x.t)
    asm.A.mult_tensor(d, y_csr)
    asm.hvp_device(x.t, d, y_mf)
    scale = float(y_csr.abs().max())
    print(f"P4 slope h=1/7: |H d - hvp d|_max / |H d|_max = {float((y_csr - y_mf).abs().max()) / scale:.3e}", flush=True)
    assert float((y_csr - y_mf).abs().max()) <= 1e-10 * scale
    H = asm.A.to_scipy()
    assert abs(H - H.T).max() <= 1e-12 * abs(H).max()
    h = 1e-6 * float(torch.linalg.vector_norm(x.t)) / float(torch.linalg.vector_norm(d))
    ep = float(asm.energy_device(x.t + h * d).item())
    em = float(asm.energy_device(x.t - h * d).item())
    gd = float(torch.dot(g.t, d))
    assert _rel((ep - em) / (2 * h), gd) <= 1e-6
    asm.cleanup()
