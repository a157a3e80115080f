"""Newton-step assembly benchmark (energy + gradient + CSR Hessian) on one MI355X.

Workload: Ginzburg-Landau, eps = 0.01 on [-1,1]^2, structured 1024x1024 grid of P1
triangles (2,097,152 elements, 1,046,529 free DOFs, 7,317,521 CSR nonzeros), initial
guess 0.5 sin(pi x) sin(pi y) + U[-1e-3, 1e-3] (seed 0).  One step = one Newton-step
assembly: energy + gradient + CSR Hessian from one fused staged-patch kernel plus the energy
finalisation kernel (Hessian values FP32, element math, energy and gradient FP64).  The
kernel reads each triangle's geometry (dvx, dvy, vol) from per-element arrays
(``--element-geometry nodes`` evaluates it from the vertex coordinates instead, bit for bit the
same values).  Inputs are resident in HBM before timing starts.

Beside the headline the line carries: ``headline_fp64`` (the same workload with FP64 Hessian
values), ``headline_geometry_nodes`` (the FP32 workload with the geometry evaluated from the
vertex coordinates), ``secondary`` (assembly and SpMV of the P4 Mohr-Coulomb slope at 1.57M DOFs and of the
Neo-Hookean tet beams at 61,440 and 491,520 elements), ``time_to_tol`` (the metric's second
half: Newton time to tolerance on the Neo-Hookean load path, the reference's solver profile in
FP64 against FP32 Hessians from the learned warm start, network trained in the same run),
``time_to_tol_headline`` (the same on the headline workload: trust-region Newton with FP64 against FP32
Hessians, no warm start), ``roofline`` and ``cpu_baseline``.

Prints one JSON line (see the field comments below).  ``--gpus N`` with N > 1 exits: this
build targets one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024, help="cells per side")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--element-geometry", default="arrays", choices=["auto", "nodes", "arrays"],
                    help="P1 element geometry from the vertex coordinates (auto: when they reproduce the "
                         "geometry arrays bit for bit) or from per-element arrays")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=256)
    ap.add_argument("--cpu-budget-s", type=float, default=15.0)
    ap.add_argument("--no-secondary", action="store_true", help="skip the P4 slope and tet beam assembly lines")
    ap.add_argument("--no-fp64-line", action="store_true", help="skip the FP64-Hessian line of the headline workload")
    ap.add_argument("--no-time-to-tol", action="store_true",
                    help="skip the times to tolerance (Neo-Hookean load path: plain FP64 vs mixed + learned; "
                         "the headline workload: FP64 vs FP32 Hessians)")
    ap.add_argument(
        "--copy-calib-mib", type=int, default=0,
        help="before warm-up, run 3 device copies of this many MiB (counter calibration reference)",
    )
    return ap.parse_args()


def cpu_baseline(n_cells: int, budget_s: float) -> dict:
    """Oracle (numpy FP64, one thread) energy + gradient + Hessian assembly on a bounded sample."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import numpy as np

    from fenics_nonlinear_energies_amd.problems.ginzburg_landau import build_gl_problem, sine_noise_init
    from oracle import assembly

    params, _ = build_gl_problem(n_cells)
    u = sine_noise_init(params, 0)
    times = []
    t_end = time.perf_counter() + budget_s
    while True:
        t0 = time.perf_counter()
        assembly.assemble("gl", params, u)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end or len(times) >= 20:
            break
    med = float(np.median(times))
    ne = params["elems"].shape[0]
    return {
        "value": ne / med,
        "unit": "elem/s",
        "cores": 1,
        "kind": "port",
        "sample": f"GL {n_cells}x{n_cells} P1 ({ne} elements), numpy FP64 element energy+gradient+Hessian "
        f"with np.add.at scatter into CSR, median of {len(times)} repeats",
    }


def cpu_baseline_omp(n_cells: int, budget_s: float) -> dict | None:
    """The C restatement of the same assembly (oracle/fne_oracle.c) over all host threads
    (OpenMP), on the full workload; None when the oracle library is not built."""
    import numpy as np

    from fenics_nonlinear_energies_amd.problems.ginzburg_landau import build_gl_problem, sine_noise_init

    try:
        from oracle import c_assembly

        c_assembly.lib()
    except (ImportError, OSError):
        return None
    params, _ = build_gl_problem(n_cells)
    u = sine_noise_init(params, 0)
    asm = c_assembly.P1Assembly("gl", params)
    asm.run(u)  # warm-up (first touch of the output and scratch pages)
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < 20:
        t0 = time.perf_counter()
        asm.run(u)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end and len(times) >= 5:
            break
    ne = params["elems"].shape[0]
    return {
        "value": ne / float(np.median(times)),
        "unit": "elem/s",
        "cores": c_assembly.threads(),
        "kind": "port",
        "sample": f"the full GL {n_cells}x{n_cells} P1 workload ({ne} elements): C/OpenMP restatement, FP64 element "
        f"energy+gradient+Hessian then a per-row sum into the CSR, {c_assembly.threads()} threads, median of "
        f"{len(times)} repeats",
    }


def _time_ms(fn, reps=20):
    import torch

    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(torch.cuda.current_stream())
    for _ in range(reps):
        fn()
    b.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def p1_algorithmic_bytes(asm, n_nodes: int, s_h: int) -> int:
    """Bytes one fused P1 assembly must move: connectivity (3 x 4 B per element), the element
    geometry (7 FP64 per element, or 16 B of vertex coordinates per node when the kernels
    evaluate it from the coordinates), nodal state (8 B per node), gradient write (8 B per free
    DOF) and CSR values write (s_h per nonzero)."""
    ne, nnz, nf = asm.n_elem, asm.layout.nnz, asm.layout.n_free
    geo = 16 * n_nodes if asm.node_geometry else 8 * 7 * ne
    return int(ne * 4 * 3 + geo + 8 * n_nodes + 8 * nf + s_h * nnz)


def headline_variant(n_cells: int, precision: str, element_geometry: str, steps: int = 50) -> dict:
    """The headline workload in another form (FP64 Hessian values, the reference's precision;
    or geometry read from per-element arrays): elem/s and the HBM fraction of the same fused
    kernel, timed like the headline."""
    import torch

    from fenics_nonlinear_energies_amd.problems.ginzburg_landau import (
        GLReorderedElementAssembler,
        build_gl_problem,
        sine_noise_init,
    )

    params, adj = build_gl_problem(n_cells)
    asm = GLReorderedElementAssembler(params, None, adj, hessian_precision=precision,
                                      element_geometry=element_geometry)
    x = asm.create_vec(sine_noise_init(params, 0)[asm.layout.perm])
    g = x.duplicate()
    ms = _time_ms(lambda: asm.energy_grad_hessian_device(x.t, g.t), reps=steps)
    ne = asm.n_elem
    b_asm = p1_algorithmic_bytes(asm, params["nodes"].shape[0], 4 if precision == "fp32" else 8)
    geo = "vertex coordinates" if asm.node_geometry else "per-element geometry arrays"
    asm.cleanup()
    torch.cuda.synchronize()
    return {"workload": f"Ginzburg-Landau {n_cells}x{n_cells} P1, fused energy + gradient + Hessian, "
                        f"Hessian {precision}, geometry from {geo}",
            "elem_per_s": ne / (ms * 1e-3), "assembly_ms": ms, "algorithmic_bytes": int(b_asm),
            "achieved_GBs": b_asm / (ms * 1e-3) / 1e9, "frac": b_asm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def time_to_tol_beam() -> dict:
    """Newton time to tolerance on the Neo-Hookean beam load path (160x8x8 hexes x 6 tets, 38,637
    free DOFs, 24 rotation steps to 4 pi, every step to ||g|| <= 2e-4): the reference's solver
    profile in FP64 from the previous state against FP32 Hessians from the learned warm start.
    The network (10-128x4-3, BF16 MFMA kernels) is trained in this run on 8 seeded runs of the
    80x4x4 beam (E ~ U[1e8, 4e8], nu ~ U[0.25, 0.4]); the evaluation material is unseen."""
    import numpy as np

    from fenics_nonlinear_energies_amd.experiments import he_warmstart_study

    r = np.random.default_rng(0)
    mats = [(float(r.uniform(1e8, 4e8)), float(r.uniform(0.25, 0.4))) for _ in range(8)]
    nk = dict(tolf=0.0, tolg=2e-4, tolg_rel=0.0, tolx_rel=0.0, tolx_abs=0.0, maxit=100, require_all_convergence=False)
    t0 = time.perf_counter()
    out = he_warmstart_study(train_materials=mats, eval_material=(2e8, 0.3), newton_kwargs=nk, log=lambda *a: None)
    total = time.perf_counter() - t0
    v = out["variants"]
    ref, lrn = v["reference_fp64"], v["learned_fp32"]
    return {
        "workload": "Neo-Hookean beam 160x8x8 hexes x 6 tets, 24 rotation steps to 4 pi, Newton to ||g|| <= 2e-4",
        "plain_fp64_s": ref["wall_s"],
        "mixed_learned_s": lrn["wall_s"],
        "speedup": ref["wall_s"] / lrn["wall_s"],
        "final_energy_rel_diff": abs(lrn["final_energy"] - ref["final_energy"]) / abs(ref["final_energy"]),
        "variants": {k: {"wall_s": x["wall_s"], "newton_its": x["newton_its"], "linear_its": x["linear_its"],
                         "all_converged": x["all_converged"], "final_energy": x["final_energy"]}
                     for k, x in v.items()},
        "training": {"samples": out["train_samples"], "loss_first": out["loss_first"], "loss_last": out["loss_last"],
                     "train_s": out["train_time_s"], "data_runs_s": out["data_gen_time_s"]},
        "study_total_s": total,
    }


def time_to_tol_gl(n_cells: int) -> dict:
    """Newton time to tolerance on the headline workload itself (Ginzburg-Landau n x n, its
    sine-plus-noise start, seed 0): trust-region Newton with Steihaug CG and the multigrid, FP64
    Hessians with the multigrid set up at every Newton iteration (the plain FP64 form, the
    reference's set-up policy) against FP32 Hessian values and Krylov products (energy, gradient
    and residual FP64), with the same policy (``speedup``: precision alone) and with the
    multigrid kept across Newton iterations, rebuilt after a Krylov solve of more than 5
    iterations (``speedup_mixed_form``; ``speedup_policy`` is the same policy in FP64), all after
    a warm-up on a small grid."""
    from fenics_nonlinear_energies_amd.experiments import gl_precision_study

    t0 = time.perf_counter()
    out = gl_precision_study(n=n_cells, seed=0, log=lambda *a: None)
    v64, v32 = out["variants"]["fp64"], out["variants"]["fp32"]
    return {
        "workload": f"Ginzburg-Landau {n_cells}x{n_cells} P1 ({out['n_free']} free DOFs), trust-region Newton + "
                    "Steihaug CG + multigrid, to |g| <= 1e-6",
        "plain_fp64_s": v64["wall_s"],
        "mixed_fp32_s": v32["wall_s"],
        "speedup": out["speedup"],
        "final_energy_rel_diff": out["energy_rel_diff"],
        "mixed_fp32_reuse_s": out["variants"]["fp32_reuse"]["wall_s"],
        "speedup_mixed_form": out["speedup_mixed_form"],
        "speedup_policy": out["speedup_policy"],
        "final_energy_rel_diff_mixed_form": out["energy_rel_diff_mixed_form"],
        "variants": {k: {x: v[x] for x in ("wall_s", "newton_its", "linear_its", "pc_setups", "energy", "message")}
                     for k, v in out["variants"].items()},
        "study_total_s": time.perf_counter() - t0,
    }


def copy_ceiling_GBs(mib: int = 1024, reps: int = 10) -> float:
    """Bytes moved per second by a plain device-to-device copy of ``mib`` MiB (read + write),
    the practical HBM ceiling the roofline fraction is read against."""
    import torch

    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(torch.cuda.current_stream())
    for _ in range(reps):
        dst.copy_(src)
    b.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    gbs = 2 * src.numel() * 4 * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del src, dst
    return gbs


def secondary_workloads() -> list[dict]:
    """Fused Newton-step assembly and SpMV on the P4 slope (h = 1/7, 1.57M DOFs) and the tet beams
    (160x8x8 and 320x16x16 hexes x 6 tets)."""
    import numpy as np

    from fenics_nonlinear_energies_amd.problems import slope_stability as S
    from fenics_nonlinear_energies_amd.problems.hyperelasticity import HEReorderedElementAssembler, build_beam_problem

    out = []
    geo = dict(x1=20.0, x2=20.0, x3=20.0, y1=10.0, y2=10.0)
    mat = dict(E=20000.0, nu=0.49, c0=6.0, phi_deg=45.0, psi_deg=0.0, gamma=20.0)
    p, _, _ = S.build_slope_problem(1.0 / 7.0, degree=4, geometry=geo, material=mat, with_adjacency=False)
    asm = S.SlopeStabilityReorderedElementAssembler(p, None, None, hessian_precision="fp32")
    X = p["nodes"]
    v = np.zeros(2 * X.shape[0])
    v[0::2] = 1e-3 * X[:, 1]
    v[1::2] = -5e-4 * X[:, 1]
    x = asm.create_vec(v[p["freedofs"]][asm.layout.perm])
    g, y = x.duplicate(), x.duplicate()
    ms = _time_ms(lambda: asm.energy_grad_hessian_device(x.t, g.t))
    sp = _time_ms(lambda: asm.A.mult(x, y))
    hv = x.t.clone()
    mf = _time_ms(lambda: asm.hvp_device(x.t, y.t, hv))  # matrix-free H v at the same state (FP64)
    ne, nnz, nf = asm.n_elem, asm.layout.nnz, asm.layout.n_free
    out.append({"workload": "Mohr-Coulomb P4 slope 60x20 m, h=1/7 (plastic state)", "n_elem": ne, "n_free": nf,
                "nnz": int(nnz), "assembly_ms": ms, "elem_per_s": ne / (ms * 1e-3), "spmv_ms": sp,
                "spmv_GBs": (8 * nnz + 20 * nf) / (sp * 1e-3) / 1e9,
                "mfma_TFLOPs": ne * 112.9e3 / (ms * 1e-3) / 1e12, "matfree_hvp_ms": mf})
    asm.cleanup()
    del asm
    for dims in ((160, 8, 8), (320, 16, 16)):
        p, adj, u0 = build_beam_problem(*dims)
        asm = HEReorderedElementAssembler(p, None, adj, hessian_precision="fp32")
        x = asm.create_vec((u0 + 1e-4 * np.random.default_rng(0).standard_normal(u0.size))[asm.layout.perm])
        g, y = x.duplicate(), x.duplicate()
        ms = _time_ms(lambda: asm.energy_grad_hessian_device(x.t, g.t))
        sp = _time_ms(lambda: asm.A.mult(x, y))
        ne, nnz, nf = asm.n_elem, asm.layout.nnz, asm.layout.n_free
        b_asm = ne * (4 * 12 + 8 * 13) + 8 * asm.dirichlet_full.size + 8 * nf + 4 * nnz
        out.append({"workload": "Neo-Hookean beam %dx%dx%d hexes x 6 tets" % dims, "n_elem": ne, "n_free": nf,
                    "nnz": int(nnz), "assembly_ms": ms, "elem_per_s": ne / (ms * 1e-3),
                    "assembly_GBs": b_asm / (ms * 1e-3) / 1e9, "assembly_frac": b_asm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "spmv_ms": sp, "spmv_GBs": (8 * nnz + 20 * nf) / (sp * 1e-3) / 1e9})
        asm.cleanup()
        del asm
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from fenics_nonlinear_energies_amd.problems.ginzburg_landau import (
        GLReorderedElementAssembler,
        build_gl_problem,
        sine_noise_init,
    )

    torch.cuda.set_device(0)
    params, adj = build_gl_problem(args.n)
    asm = GLReorderedElementAssembler(params, None, adj, hessian_precision=args.precision,
                                      element_geometry=args.element_geometry)
    u = sine_noise_init(params, 0)
    x = asm.create_vec(u[asm.layout.perm])
    g = x.duplicate()
    y = x.duplicate()
    stream = torch.cuda.current_stream()
    if args.copy_calib_mib > 0:
        # a plain device copy of known size: HBM counters of this kernel calibrate the assembly's
        src = torch.empty(args.copy_calib_mib << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def step():
        asm.energy_grad_hessian_device(x.t, g.t)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # timed region: K steps back to back, bracketed by two HIP events on the launch stream (the
    # assembly's launch duration is their span / K: the fused kernel plus its energy
    # finalisation).  No event inside the loop: a pair of event records around every step
    # added ~7 us of stream markers per 55 us step.
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps

    # SpMV (the Krylov hot loop) on the assembled Hessian
    for _ in range(10):
        asm.A.mult(x, y)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    n_sp = 100
    for _ in range(n_sp):
        asm.A.mult(x, y)
    e.record(stream)
    torch.cuda.synchronize()
    spmv_ms = s.elapsed_time(e) / n_sp

    copy_gbs = copy_ceiling_GBs()

    ne = asm.n_elem
    nnz = asm.layout.nnz
    nf = asm.layout.n_free
    nt = params["nodes"].shape[0]
    s_h = 4 if args.precision == "fp32" else 8
    b_asm = p1_algorithmic_bytes(asm, nt, s_h)
    b_spmv = nnz * (s_h + 4) + 4 * (nf + 1) + 8 * nf + 8 * nf
    achieved = b_asm / (kern_ms * 1e-3) / 1e9
    traffic = None
    # HBM bytes per launch of the assembly kernel from the committed FETCH_SIZE/WRITE_SIZE summary
    geo_tag = "_nodegeo" if asm.node_geometry else ""
    tfile = ROOT / "profiles" / f"p1_staged_gl{args.n}_{args.precision}{geo_tag}_traffic.md"
    if tfile.exists():
        for line in tfile.read_text().splitlines():
            if line.startswith("hbm_bytes_per_launch:"):
                traffic = int(float(line.split(":")[1]))
    ms_per_step = wall / args.steps * 1e3
    out = {
        "metric": "elem/s for energy+grad+Hessian CSR assembly",
        "value": ne / (wall / args.steps),
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 element math, f32 Hessian values",
        "data": "synthetic",
        "config": {
            "workload": f"Ginzburg-Landau eps=0.01, {args.n}x{args.n} P1 grid, one Newton-step assembly "
            f"(fused energy + gradient + Hessian), Hessian {args.precision}",
            "element_geometry": "vertex coordinates" if asm.node_geometry else "per-element arrays",
            "n_elem": ne,
            "n_free": nf,
            "nnz": int(nnz),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "p1_staged_kernel (fused energy + gradient + Hessian) + finalize_pairs_kernel",
            "algorithmic_bytes_per_launch": int(b_asm),
            "avg_launch_ms": kern_ms,
            # practical ceiling: a plain 1 GiB device-to-device copy (read + write) on this GPU
            "copy_GBs": copy_gbs,
            "frac_of_copy": achieved / copy_gbs,
            "traffic_GBs": None if traffic is None else traffic / (kern_ms * 1e-3) / 1e9,
        },
        "spmv": {
            "avg_ms": spmv_ms,
            "achieved_GBs": b_spmv / (spmv_ms * 1e-3) / 1e9,
            "frac": b_spmv / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "frac_of_copy": b_spmv / (spmv_ms * 1e-3) / 1e9 / copy_gbs,
            "algorithmic_bytes": int(b_spmv),
        },
    }
    node_geo = bool(asm.node_geometry)
    asm.cleanup()
    if not args.no_fp64_line and args.precision == "fp32":
        out["headline_fp64"] = headline_variant(args.n, "fp64", args.element_geometry)
        other = "arrays" if node_geo else "nodes"
        out[f"headline_geometry_{other}"] = headline_variant(args.n, args.precision, other)
    if not args.no_secondary:
        out["secondary"] = secondary_workloads()
    if not args.no_time_to_tol:
        out["time_to_tol"] = time_to_tol_beam()
        out["time_to_tol_headline"] = time_to_tol_gl(args.n)
    if not args.no_cpu_baseline:
        single = cpu_baseline(args.cpu_sample_n, args.cpu_budget_s)
        omp = cpu_baseline_omp(args.n, args.cpu_budget_s)
        out["cpu_baseline"] = omp if omp is not None else single
        if omp is not None:
            out["cpu_baseline_single_thread"] = single
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
