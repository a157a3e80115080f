"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

A plain numpy (and C, see ``fne_oracle.c``) restatement of the reference's hot-path
algorithm: element energies with hand-derived gradients/Hessians, the np.add.at scatter
into the owned CSR pattern, the energy/gradient callbacks and the Newton / line-search /
trust-region control flow.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or the timed
CPU baseline.  The product package (``fenics_nonlinear_energies_amd``) never imports it.

Each function cites the reference file:line it follows.  How each function is pinned
(golden vectors, closed forms, finite differences, autodiff) is listed in ``PINNING`` below
and enforced by tests/test_oracle_*.py.
"""

PINNING = {
    "elements.plaplace": "FD + torch autodiff twin + p=2 Laplacian closed form + reference energies (HDF5 meshes, docs tables)",
    "elements.ginzburg_landau": "FD + torch autodiff twin + eps-only Laplacian closed form + reference energies",
    "elements.neohookean": "FD + torch autodiff twin + F=I linear-elasticity closed form + reference energies",
    "elements.mohr_coulomb": "FD + torch autodiff twin + elastic-branch B^T C B closed form + yield-surface invariant",
    "assembly": "brute-force dense double loop and scipy COO->CSR duplicate summation",
    "minimizers": "quadratic/1-D known answers; golden-section exactness on parabolas",
    "solvers.cg/gmres/stcg/golden_section": "exact solves, k-distinct-eigenvalue termination, steepest-descent first step, Steihaug's monotone path, boundary and negative-curvature stops, golden-section evaluation count (tests/test_oracle_solvers.py)",
    "davis_b": "reference golden values tests/test_slope_stability_jax.py:20-24",
    "c_assembly (fne_oracle.c)": "equal to assembly.assemble (energy, gradient, CSR Hessian to 1e-13) for p-Laplace and GL",
}
