#!/usr/bin/env python3
"""Benchmark: P2-tet isotropic elasticity assembly on one MI355X.

Workload: lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u, vector
FEM_PK(3,2), IM_TETRAHEDRON(5), on the jittered unit cube split into 64^3
cubes x 6 tetrahedra (1,572,864 tets, 6,440,067 dofs), per-element lognormal
lambda, mu.  One step = one order-2 assembly (geometry kernel + contraction /
scatter kernel) into the prebuilt pattern, inputs resident in HBM.

Prints ONE JSON line (rank 0).  Extra keys: roofline (dominant kernel),
cpu_baseline (the oracle port on host cores over a bounded sample), forms
(every contraction form / precision with its error against FP64).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "P2-tet elasticity assembly elems/s + HBM GB/s; FP32 vs FP64 form at stated error"
HBM_PEAK_GBS = 8000.0


def bytes_per_element(n_nodes, n_geo, n_coef, n_loc, n_state=0):
    """Algorithmic bytes per element: connectivity + geometry + coefficients +
    an FP64 read, an FP64 write and a 4-byte CSR position per element-matrix
    entry."""
    return 4 * n_nodes + 24 * n_geo + 8 * n_coef + 8 * n_state + 20 * n_loc * n_loc


CONFIGS = {
    # name: (n, fem, qdim, im, term, n_nodes, n_coef, reference nnz at the default n)
    "c1": (20, "FEM_PK(3,1)", 1, "IM_TETRAHEDRON(1)", "laplace", 4, 0, 128113),
    "c2": (64, "FEM_PK(3,2)", 3, "IM_TETRAHEDRON(5)", "elast", 10, 2, 544886097),
    "c5": (32, "FEM_PK(3,3)", 3, "IM_TETRAHEDRON(8)", "elast", 20, 2, 386702445),
    "c3": (48, "FEM_QK(3,2)", 3, "IM_GAUSS_PARALLELEPIPED(3,5)", "neohookean", 27, 2, 513599625),
}
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix peak
# HBM bytes per assembly of the dominant kernel(s), measured with rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch, summed over one assembly's
# launches) at each config's default size; FETCH_SIZE is scaled by 2 and
# WRITE_SIZE by 1, the factors a plain 1 GiB device copy gives
# (bench.py --copy-calibration).  Every form of a config has the same memory
# pattern (FP64 values and accumulation), so one figure serves all forms.
PMC_TRAFFIC_KIB = {  # config: (kernel, FETCH_SIZE KiB, WRITE_SIZE KiB)
    "c2": ("k_assemble_patch", 455101, 4521372),
    "c5": ("k_assemble_patch", 180081, 3182486),
    "c3": ("k_nh_element", 55194, 6042268),
}
FETCH_FACTOR, WRITE_FACTOR = 2.0, 1.0


def pmc_traffic_bytes(config, n, n0):
    if config not in PMC_TRAFFIC_KIB or n != n0:
        return None
    _, f, w = PMC_TRAFFIC_KIB[config]
    return (FETCH_FACTOR * f + WRITE_FACTOR * w) * 1024.0
WORKLOAD = {
    "c1": "P1-tet Poisson Grad_u.Grad_Test_u",
    "c2": "P2-tet isotropic elasticity lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u",
    "c5": "P3-tet isotropic elasticity lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u",
    "c3": "Q2-hex compressible Neo-Hookean (Bonet) tangent + residual, lambda=10, mu=1, "
          "u = sum of 8 sine modes (max |u| = 0.05)",
}


def run_gpu(args):
    import torch

    import getfem_amd as ga
    from getfem_amd import meshes

    n0, fem, q, im, term, n_nodes, n_coef, ref_nnz = CONFIGS[args.config]
    n = args.n or n0
    t0 = time.perf_counter()
    hexes = term == "neohookean"
    pts, cvx = meshes.regular_unit_mesh(n, "GT_QK(3,1)" if hexes else "GT_PK(3,1)")
    if hexes:
        lam, mu = 10.0, 1.0
    else:
        pts = meshes.jitter(pts, n)
        lam, mu = meshes.lognormal_lame(cvx.shape[0])
    t_inputs = time.perf_counter() - t0
    n_cv = cvx.shape[0]
    print("inputs: n=%d, %d tets (%.1f s)" % (n, n_cv, t_inputs), file=sys.stderr, flush=True)

    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ctx = ga.Context(0)
    ctx.set_stream(stream.cuda_stream)
    mesh = ga.Mesh(ctx, "GT_QK(3,1)" if hexes else "GT_PK(3,1)", pts, cvx)
    t0 = time.perf_counter()
    mf = ga.MeshFem(mesh, fem, q)
    t_dofs = time.perf_counter() - t0
    mim = ga.MeshIm(mesh, im)
    state = None
    if term == "laplace":
        plan = ga.Plan(mf, mim, ga.TERM_LAPLACE, [1.0])
    elif hexes:
        plan = ga.Plan(mf, mim, ga.TERM_NEOHOOKEAN, [lam, mu])
        state = meshes.sine_modes()(mf.basic_dof_nodes(), 0.05).ravel()
        plan.set_state(state)
    else:
        plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, [lam, mu])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nnz = plan.build_pattern()
    torch.cuda.synchronize()
    t_pattern = time.perf_counter() - t0
    ndof = mf.nbdof()
    print("dofs %d (%.1f s host), pattern nnz %d (%.3f s GPU)" % (ndof, t_dofs, nnz, t_pattern),
          file=sys.stderr, flush=True)

    n_loc = n_nodes * q
    if hexes:  # 8 geometric nodes, state gathered, residual assembled
        b_elem = bytes_per_element(n_nodes, 8, n_coef, n_loc, n_state=n_loc) + 16 * n_loc
    else:
        b_elem = bytes_per_element(n_nodes, 4, n_coef, n_loc)
    b_patch = 4 * n_nodes + 20 * n_loc * n_loc  # the contraction/scatter kernel's share
    variants = [("fp64_factorized", ga.FORM_FACTORIZED, ga.PREC_FP64),
                ("fp64_perqp", ga.FORM_PERQP, ga.PREC_FP64),
                ("fp32_factorized", ga.FORM_FACTORIZED, ga.PREC_FP32),
                ("fp32_perqp", ga.FORM_PERQP, ga.PREC_FP32)]
    if term != "laplace" and not hexes and n_nodes in (10, 20):
        variants.append(("bf16x2_element_batched", ga.FORM_FACTORIZED, ga.PREC_BF16X2))
    if hexes:
        variants = [("fp64_element_batched", ga.FORM_PERQP, ga.PREC_FP64),
                    ("fp32_element_batched", ga.FORM_PERQP, ga.PREC_FP32)]
    if args.form != "all":
        variants = [v for v in variants if v[0] == args.form]
    forms = {}
    ref_vals = None
    for name, form, prec in variants:
        plan.set_form(form, prec)
        for _ in range(args.warmup):
            plan.assemble()
        torch.cuda.synchronize()
        plan.set_timing(True)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            plan.assemble()
        e1.record(stream)
        torch.cuda.synchronize()
        ms_step = e0.elapsed_time(e1) / args.steps
        calls, ms_geo, ms_asm = plan.timing_summary()
        plan.set_timing(False)
        ms_geo /= calls
        ms_asm /= calls
        _, _, vals = plan.torch_csr(dev)
        if ref_vals is None and prec == ga.PREC_FP64:
            ref_vals = vals.clone()
        err = None
        if ref_vals is not None:
            err = float(torch.linalg.vector_norm(vals - ref_vals) / torch.linalg.vector_norm(ref_vals))
        del vals
        forms[name] = {
            "ms_per_step": ms_step,
            "elems_per_s": n_cv / (ms_step * 1e-3),
            "gbs_algorithmic": b_elem * n_cv / (ms_step * 1e-3) / 1e9,
            "ms_geometry_kernel": ms_geo,
            "ms_assemble_kernel": ms_asm,
            "rel_frobenius_err_vs_fp64": err,
        }
        print(name, json.dumps(forms[name]), file=sys.stderr, flush=True)
    # integer checks of the pattern at full size
    ip, ix, _ = plan.torch_csr(dev)
    rowlen = ip[1:] - ip[:-1]
    pattern = {
        "nnz": int(nnz),
        "nnz_equals_reference": (int(nnz) == ref_nnz) if n == n0 else None,
        "ndof": int(ndof),
        "build_s": t_pattern,
        "dof_enumeration_s_host": t_dofs,
        "max_row_len": int(rowlen.max().item()),
    }
    del ip, ix
    head = forms[variants[0][0]]
    # compulsory bytes of one assembly: CSR values + column indices + row
    # pointers + points + connectivity + per-element coefficients
    compulsory = (12 * int(nnz) + 8 * (int(ndof) + 1) + 24 * pts.shape[0] + 4 * cvx.size
                  + 8 * n_coef * n_cv)
    out = {
        "metric": METRIC,
        "value": head["elems_per_s"],
        "unit": "elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if "fp64" in variants[0][0] else "f32",
        "data": "synthetic (seeded jittered mesh, lognormal lambda/mu)",
        "config": {
            "workload": ("%s, %d^3 = %d hexes, %d dofs, %s; step = element matrices and residuals + "
                         "node gathers into the prebuilt CSR" % (WORKLOAD[args.config], n, n_cv, ndof, im))
                        if hexes else
                        ("%s, %d^3x6 = %d tets, %d dofs, %s, per-element lambda/mu; step = geometry + "
                         "contraction/scatter into the prebuilt CSR" % (WORKLOAD[args.config], n, n_cv,
                                                                        ndof, im)),
            "name": args.config,
            "form": variants[0][0],
            "elements": n_cv,
            "nnz": int(nnz),
        },
        "hbm_gbs_algorithmic": head["gbs_algorithmic"],
        "roofline": {
            "bound": "hbm",
            "achieved": b_patch * n_cv / (head["ms_assemble_kernel"] * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": b_patch * n_cv / (head["ms_assemble_kernel"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": pmc_traffic_bytes(args.config, n, n0),
            "kernel": "k_assemble_patch",
            "bytes_per_element": b_patch,
            "kernel_ms": head["ms_assemble_kernel"],
            "launches_per_assembly": "one packed launch in node order plus one per size class of large patches, concurrent; kernel_ms is their span",
            "compulsory_bytes": compulsory,
            "compulsory_frac": compulsory / (head["ms_assemble_kernel"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "forms": forms,
        "pattern": pattern,
    }
    if not hexes:
        out["mfma"] = {"used": False,
                       "note": "3x3 pair blocks on the vector ALUs (FP64 matrix rate = vector rate on MI355X)"}
        bdb = forms.get("bf16x2_element_batched")
        if bdb:
            # element matrices sum_q sign(w_q) F_q^T F_q (F = sqrt(|w_q||J|) S B_q, S S = D)
            # on the matrix cores: ceil(nq/4) k-steps of 32 (4 points x 8 Voigt rows, 2 of
            # them zero) x nt (nt + 1) / 2 upper 16 x 16 tiles (nt = n_loc/16 rounded up;
            # the lower ones are mirrored) x 3 bf16 products of v_mfma_f32_16x16x32_bf16
            # (16384 FLOP each)
            nq = mim.points()[1].size
            nt = -(-n_loc // 16)
            n_mfma = -(-nq // 4) * (nt * (nt + 1) // 2) * 3
            issued = n_mfma * 2 * 16 * 16 * 32
            t_el = bdb["ms_geometry_kernel"] * 1e-3
            out["mfma"] = {
                "form": "bf16x2_element_batched", "kernel": "k_elast_bdb_mfma",
                "instruction": "v_mfma_f32_16x16x32_bf16", "mfma_per_element": n_mfma,
                "issued_tflops": issued * n_cv / t_el / 1e12, "peak_tflops": BF16_PEAK_TFLOPS,
                "frac": issued * n_cv / t_el / 1e12 / BF16_PEAK_TFLOPS,
                "useful_flop_per_element": 2 * nq * (6 * n_loc * n_loc + 36 * n_loc),
                "useful_tflops_equivalent": 2 * nq * (6 * n_loc * n_loc + 36 * n_loc) * n_cv / t_el / 1e12,
                "ms_records_plus_element_kernel": bdb["ms_geometry_kernel"],
                "ms_gather": bdb["ms_assemble_kernel"],
                "note": "the FP64 node-patch form (36 kFLOP-class factorized pair blocks) stays the "
                        "headline; this form buffers n_loc^2 FP32 per element and gathers them",
            }
    if hexes:
        # the element kernel (per-point law + 81x81 contraction) is FP64-compute bound;
        # FLOPs per element of the per-point B^T D B model: 2 nq (n_loc^2 d + n_loc d^2), d = 9
        flop = 2 * 27 * (n_loc * n_loc * 9 + n_loc * 81)
        t_el = head["ms_geometry_kernel"] * 1e-3
        out["roofline"] = {
            "bound": "mfma", "achieved": flop * n_cv / t_el / 1e12, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": flop * n_cv / t_el / 1e12 / FP64_PEAK_TFLOPS,
            "traffic": pmc_traffic_bytes(args.config, n, n0),
            "kernel": "k_nh_element", "flop_per_element": flop, "kernel_ms": head["ms_geometry_kernel"],
            "gather_kernel": {"kernel": "k_gather_emat", "ms": head["ms_assemble_kernel"],
                              "achieved_gbs": b_patch * n_cv / (head["ms_assemble_kernel"] * 1e-3) / 1e9},
        }
        out["dtype"] = "f64"
        f32 = forms.get("fp32_element_batched")
        if f32:
            # the FP32 form's G^T T product on the matrix cores: per element 4 k-steps x 18
            # (a, column tile) pairs x 2 row tiles x 6 bf16 products (bf16x3 split) of
            # v_mfma_f32_16x16x32_bf16 (16384 FLOP each), padding included
            n_mfma = 4 * 18 * 2 * 6
            issued = n_mfma * 2 * 16 * 16 * 32
            t32 = f32["ms_geometry_kernel"] * 1e-3
            out["mfma"] = {
                "form": "fp32_element_batched", "kernel": "k_nh_element",
                "instruction": "v_mfma_f32_16x16x32_bf16", "mfma_per_element": n_mfma,
                "issued_tflops": issued * n_cv / t32 / 1e12, "peak_tflops": BF16_PEAK_TFLOPS,
                "frac": issued * n_cv / t32 / 1e12 / BF16_PEAK_TFLOPS,
                "useful_fp32_flop_per_element": 2 * 3 * 27 * 81 * 81,
                "kernel_ms": f32["ms_geometry_kernel"],
            }
    return out, (pts, cvx, lam, mu, fem, q, im, term, state)


C4_NU = 0.3  # Poisson ratio of the SIMP problem (lambda, mu of E = 1)


def run_c4(args):
    """SIMP compliance J = u^T K(rho) u on P1 tets (32^3x6) with its three
    gradients (dJ/drho, dJ/dX, dJ/du) in one device call per step."""
    import torch

    import getfem_amd as ga
    from getfem_amd import meshes
    from getfem_amd.torch_ops import Compliance

    n = args.n or 32
    lam = C4_NU / ((1 + C4_NU) * (1 - 2 * C4_NU))
    mu = 1 / (2 * (1 + C4_NU))
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    dev = torch.device("cuda:0")
    ctx = ga.Context(0)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    op = Compliance(mf, mim, lam, mu, 1e-3, 3.0)
    rho = torch.tensor(meshes.simp_density(cvx.shape[0]), device=dev)
    u = torch.tensor(meshes.random_state(mf.nbdof()), device=dev)
    X = torch.tensor(pts, device=dev)
    n_cv = cvx.shape[0]
    outs = (torch.empty((), dtype=torch.float64, device=dev), torch.empty_like(rho),
            torch.empty_like(X), torch.empty_like(u))
    # the call is kernels only: capture it once in a graph and replay it (the
    # launch cost of one call is above its kernel time at this size)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(max(1, args.warmup)):
            op.eval(rho, X, u, out=outs)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        op.eval(rho, X, u, out=outs)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    J = outs[0]
    # the same step without the graph, for reference
    e0.record()
    for _ in range(args.steps):
        op.eval(rho, X, u, out=outs)
    e1.record()
    torch.cuda.synchronize()
    ms_eager = e0.elapsed_time(e1) / args.steps
    b_elem = 416  # 4*4 + 24*4 + 8 + 8*12 + 8 + 16*12 bytes per element (gradient pass)
    out = {
        "metric": METRIC,
        "value": n_cv / (ms * 1e-3),
        "unit": "elems/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered mesh, rho ~ U(0.1,1), u ~ N(0,1))",
        "config": {"workload": "P1-tet SIMP compliance J = u^T K(rho) u with dJ/drho, dJ/dX, dJ/du, "
                               "%d^3x6 = %d tets, E = 1e-3 + rho^3 (1 - 1e-3), nu = 0.3; step = one device "
                               "call (element kernel + one launch for the two node gathers and the J sum), replayed from a HIP graph"
                               % (n, n_cv),
                   "name": "c4", "elements": n_cv},
        "ms_per_step_without_graph": ms_eager,
        "J": float(J),
        "roofline": {"bound": "hbm", "achieved": b_elem * n_cv / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": b_elem * n_cv / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_compliance_elem + k_compliance_finish", "bytes_per_element": b_elem,
                     "kernel_ms": ms,
                     "note": "82 MB per step: cache-resident on MI355X (256 MB last-level cache); the step "
                             "is launch/latency bound"},
    }
    return out, (pts, cvx, mf.basic_dof_table(), mf.nbdof(), lam, mu, rho.cpu().numpy(), u.cpu().numpy())


def cpu_baseline_c4(inputs):
    """The oracle's compliance with its analytic gradients, repeated over the
    whole mesh for about 5 s (one thread)."""
    import oracle
    pts, cvx, dofs, ndof, lam, mu, rho, u = inputs
    prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, 1, 3, dofs, ndof, "IM_TETRAHEDRON(1)",
                          oracle.TERM_ISO_ELAST, scal_a=lam, scal_b=mu, U=u)
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 5.0:
        prob.compliance(rho, 1e-3, 3.0, grads=True)
        reps += 1
    dt = time.perf_counter() - t0
    return {"value": reps * cvx.shape[0] / dt, "unit": "elems/s", "cores": 1, "kind": "port",
            "sample": "%d full evaluations (J, dJ/drho, dJ/dX) of the %d-element mesh, %.1f s" % (reps, cvx.shape[0], dt)}


def cpu_baseline(inputs, threads):
    """The oracle (a C port of GetFEM's per-element/per-point FP64 algorithm
    with its threshold and sorted-column insertion) on a bounded sample."""
    import oracle
    pts, cvx, lam, mu, fem, q, im, term, state = inputs
    k = int(fem[-2])
    if term == "neohookean":
        dofs, ndof = oracle.enumerate_dof(oracle.GT_QK31, pts, cvx, oracle.FEM_QK, k, q)
        prob = oracle.Problem(oracle.GT_QK31, pts, cvx, oracle.FEM_QK, k, q, dofs, ndof, im,
                              oracle.TERM_NEOHOOKEAN, scal_a=lam, scal_b=mu, U=state)
    else:
        dofs, ndof = oracle.enumerate_dof(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, q)
    if term == "neohookean":
        pass
    elif term == "laplace":
        prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, q, dofs, ndof, im,
                              oracle.TERM_LAPLACE, scal_a=1.0)
    else:
        prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, q, dofs, ndof, im,
                              oracle.TERM_ISO_ELAST, coef_a=lam, coef_b=mu)
    res = {}
    for th, m in threads:
        m = min(m, cvx.shape[0])
        t0 = time.perf_counter()
        prob.assemble_timed(0, m, th)
        dt = time.perf_counter() - t0
        res[th] = (m / dt, m, dt)
        print("cpu baseline: %d threads, %d elements, %.2f s, %.0f elems/s" % (th, m, dt, m / dt),
              file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c4"])
    ap.add_argument("--n", type=int, default=0, help="mesh subdivisions (default: the config's)")
    ap.add_argument("--form", default="all")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0, help="elements for the multi-thread CPU sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 8 << 27}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.config == "c4":
        out, inputs = run_c4(args)
        if not args.no_cpu:
            try:
                import oracle
                if not os.path.exists(oracle.LIB_PATH):
                    oracle.build()
                out["cpu_baseline"] = cpu_baseline_c4(inputs)
            except Exception as exc:  # report, never hide
                out["cpu_baseline"] = {"value": None, "error": repr(exc)}
        print(json.dumps(out), flush=True)
        return
    out, inputs = run_gpu(args)
    if not args.no_cpu:
        try:
            import oracle
            if not os.path.exists(oracle.LIB_PATH):
                oracle.build()
            cores = min(16, os.cpu_count() or 1)
            n_cv = inputs[1].shape[0]
            # samples sized for about 5 s on one core and 10 s on all cores
            # (measured rates: P2 43k/73k, P3 5.2k/19k, Q2 1.3k/7.4k elems/s)
            one, many = {"c1": (48000, 48000), "c2": (240000, 786432), "c5": (25000, 196608),
                         "c3": (6000, 73728)}[args.config]
            big = args.cpu_sample or min(n_cv, many)
            res = cpu_baseline(inputs, [(1, min(one, n_cv)), (cores, big)])
            v1 = res[1][0]
            vm, m, dt = res[cores]
            out["cpu_baseline"] = {
                "value": vm, "unit": "elems/s", "cores": cores, "kind": "port",
                "sample": "first %d of %d elements (%.1f s), OpenMP element partition with "
                          "per-thread matrices summed; 1 core: %.0f elems/s on the first %d "
                          "elements" % (m, n_cv, dt, v1, res[1][1]),
                "value_1core": v1,
            }
        except Exception as exc:  # report, never hide
            out["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
