#!/usr/bin/env python3
"""Benchmark: P2-tet isotropic elasticity assembly on one MI355X.

Workload: lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u, vector
FEM_PK(3,2), IM_TETRAHEDRON(5), on the jittered unit cube split into 64^3
cubes x 6 tetrahedra (1,572,864 tets, 6,440,067 dofs), per-element lognormal
lambda, mu.  One step = one order-2 assembly (geometry kernel + contraction /
scatter kernel) into the prebuilt pattern, inputs resident in HBM.

Prints ONE JSON line (rank 0).  Extra keys:
  roofline      the dominant kernel against the HBM peak: achieved = the
                compulsory bytes of one assembly (stored values and column
                indices, row pointers, points, connectivity, coefficients)
                over the kernel's median duration; traffic = HBM bytes per
                assembly measured in this run by two rocprofv3 PMC passes
                (FETCH_SIZE, WRITE_SIZE) calibrated on a plain device copy
  cpu_baseline  the oracle (a C port of GetFEM's per-element FP64 algorithm)
                on the host's CPU share, 1 warm-up + median of 5
  forms         every contraction form / precision with its error against FP64
"""
import argparse
import glob
import json
import os
import re
import shutil
import statistics
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "P2-tet elasticity assembly elems/s + HBM GB/s; FP32 vs FP64 form at stated error"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix peak
FP32_PEAK_TFLOPS = 157.3


def bytes_per_element(n_nodes, n_geo, n_coef, n_loc, n_state=0):
    """Element-scatter byte model: connectivity + geometry + coefficients +
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
WORKLOAD = {
    "c1": "P1-tet Poisson Grad_u.Grad_Test_u",
    "c2": "P2-tet isotropic elasticity lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u",
    "c5": "P3-tet isotropic elasticity lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u",
    "c3": "Q2-hex compressible Neo-Hookean (Bonet) tangent + residual, lambda=10, mu=1, "
          "u = sum of 8 sine modes (max |u| = 0.05)",
}
# kernels whose HBM traffic the PMC passes measure, per config
PMC_KERNELS = {"c2": "k_assemble_patch", "c5": "k_assemble_patch", "c3": "k_nh_element|k_gather"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stream_rates(nbytes=4 << 30):
    """Plain-device reference rates on this GPU, measured in the same run:
    torch's fill_ (a write stream) and copy_ (read + write) of nbytes, median
    of 10 after 3 warm-ups.  They say what HBM delivers to a simple streaming
    kernel here, beside the vendor peak."""
    import torch
    x = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    out = {}
    for name, fn, moved in (("fill", lambda: x.fill_(1.0), nbytes), ("copy", lambda: y.copy_(x), 2 * nbytes)):
        for _ in range(3):
            fn()
        ts = []
        for _ in range(10):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        out[name + "_gbs"] = moved / (statistics.median(ts) * 1e-3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    return out


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
    n_cv = cvx.shape[0]
    log("inputs: n=%d, %d elements (%.1f s)" % (n, n_cv, time.perf_counter() - t0))

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
    log("dofs %d (%.1f s host), pattern nnz %d (%.3f s GPU)" % (ndof, t_dofs, nnz, t_pattern))

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
        variants += [("fp64_element_batched_mfma", ga.FORM_FACTORIZED, ga.PREC_FP64_MFMA),
                     ("fp32_element_batched_mfma", ga.FORM_FACTORIZED, ga.PREC_FP32_MFMA),
                     ("bf16x2_element_batched", ga.FORM_FACTORIZED, ga.PREC_BF16X2)]
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
        geo, asm = plan.timing_calls()
        plan.set_timing(False)
        if args.pmc_child:
            continue
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
            "ms_geometry_kernel": float(np.median(geo)),
            "ms_assemble_kernel": float(np.median(asm)),
            "ms_assemble_kernel_min": float(np.min(asm)),
            "rel_frobenius_err_vs_fp64": err,
        }
        if hexes and not args.pmc_child:
            # a Newton step's assembly: a new state each time (scaled copies of
            # the state on the device), so every assembly re-derives the
            # thresholds and keep bits in FP64 and compares the pattern
            # (kept: unchanged) before the form's element pass and gathers
            U0 = torch.tensor(state, device=dev)
            Uk = torch.empty_like(U0)
            tk = []
            for k in range(1, 6):
                Uk.copy_(U0 * (1.0 + 0.01 * k))
                torch.cuda.synchronize()
                plan.set_state_device(Uk.data_ptr())
                t0 = time.perf_counter()
                plan.assemble()
                plan.ctx.synchronize()
                tk.append((time.perf_counter() - t0) * 1e3)
            forms[name]["ms_new_state_assembly"] = float(np.median(tk))
            forms[name]["new_state_note"] = ("set_state_device + assemble (median of 5 states): FP64 element pass "
                                             "and keep bits re-derived, pattern compared and kept, then the form's "
                                             "element pass (FP32 form only) and the gathers; host clock incl. sync")
            plan.set_state_device(U0.data_ptr())
            plan.assemble()
            plan.ctx.synchronize()
            del U0, Uk
        log(name, json.dumps(forms[name]))
    if args.pmc_child:
        return None, None
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
    k_ms = head["ms_assemble_kernel"]
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
        # HBM GB/s of the metric: the compulsory bytes of one assembly over the
        # whole step (measured traffic over kernel time: roofline.traffic_gbs)
        "hbm_gbs": compulsory / (head["ms_per_step"] * 1e-3) / 1e9,
        "roofline": {
            "bound": "hbm",
            "achieved": compulsory / (k_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": compulsory / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_assemble_patch",
            "bytes_per_launch": compulsory,
            "bytes_model": "compulsory bytes of one assembly: 12 B per stored entry (FP64 value + int32 column "
                           "index) + 8 B per row pointer + points + connectivity + per-element coefficients",
            "kernel_ms": k_ms,
            "kernel_ms_min": head["ms_assemble_kernel_min"],
            "kernel_ms_stat": "median over the timed steps (HIP events on the launch stream around the "
                              "assembly's concurrent patch launches)",
            "element_scatter_model": {
                "bytes_per_element": b_patch,
                "gb_per_assembly": b_patch * n_cv / 1e9,
                "note": "SURVEY-style element-scatter count (a read, a write and a position per "
                        "element-matrix entry); the node-patch kernel writes each CSR value once and reads "
                        "no positions, so this is not a bandwidth and is not used for frac",
            },
        },
        "forms": forms,
        "pattern": pattern,
    }
    if not args.pmc_child:
        sr = stream_rates()
        r = out["roofline"]
        r["stream_reference"] = dict(sr, note="torch fill_ / copy_ of 4 GiB on this GPU in this run: what a plain "
                                              "streaming kernel reaches here (the frac above is against the 8 TB/s "
                                              "vendor peak)")
        r["frac_of_fill_rate"] = r["achieved"] / sr["fill_gbs"]
        log("stream reference", json.dumps(sr))
    if not hexes:
        # The element-batched B^T D B forms on the matrix cores: E = sum_q sign(w_q) F_q^T F_q
        # (F = sqrt(|w_q||J|) S B_q, S S = D), upper 16 x 16 tiles only (nt = n_loc/16 rounded
        # up, nt (nt + 1) / 2 tiles; the lower ones are mirrored).
        #   bf16x2: k-steps of 32 (4 points x 8 Voigt rows, 2 of them zero), 3 bf16 products of
        #           v_mfma_f32_16x16x32_bf16 (16384 FLOP each) per tile and step
        #   fp64 / fp32: k-steps of 4 over the 6 nq Voigt rows, one v_mfma_f64_16x16x4_f64 /
        #           v_mfma_f32_16x16x4_f32 (2048 FLOP each) per tile and step
        nq = mim.points()[1].size
        nt = -(-n_loc // 16)
        tiles = nt * (nt + 1) // 2
        useful = 2 * nq * (6 * n_loc * n_loc + 36 * n_loc)
        spec = {"bf16x2_element_batched": ("k_elast_bdb_mfma", "v_mfma_f32_16x16x32_bf16",
                                           -(-nq // 4) * tiles * 3, 2 * 16 * 16 * 32, BF16_PEAK_TFLOPS),
                "fp64_element_batched_mfma": ("k_elast_bdb_mfma_hp<double>", "v_mfma_f64_16x16x4_f64",
                                              -(-6 * nq // 4) * tiles, 2 * 16 * 16 * 4, FP64_PEAK_TFLOPS),
                "fp32_element_batched_mfma": ("k_elast_bdb_mfma_hp<float>", "v_mfma_f32_16x16x4_f32",
                                              -(-6 * nq // 4) * tiles, 2 * 16 * 16 * 4, FP32_PEAK_TFLOPS)}
        mf = {}
        for fname, (kern, instr, n_mfma, flop_per, peak) in spec.items():
            f = forms.get(fname)
            if not f:
                continue
            t_el = f["ms_geometry_kernel"] * 1e-3  # records + element kernel
            issued = n_mfma * flop_per * n_cv / t_el / 1e12
            mf[fname] = {"kernel": kern, "instruction": instr, "mfma_per_element": n_mfma,
                         "issued_tflops": issued, "peak_tflops": peak, "frac": issued / peak,
                         "useful_flop_per_element": useful,
                         "useful_tflops_equivalent": useful * n_cv / t_el / 1e12,
                         "elems_per_s": f["elems_per_s"], "rel_frobenius_err_vs_fp64": f["rel_frobenius_err_vs_fp64"],
                         "ms_records_plus_element_kernel": f["ms_geometry_kernel"],
                         "ms_gather": f["ms_assemble_kernel"]}
        out["mfma"] = mf if mf else {
            "used": False, "note": "3x3 pair blocks on the vector ALUs (FP64 matrix rate = vector rate on MI355X)"}
        if mf:
            out["mfma_note"] = ("the FP64 node-patch form (factorized 3x3 pair blocks on the vector ALUs) stays the "
                                "headline; the element-batched forms buffer n_loc^2 values per element and gather them")
    if hexes:
        # The step is two launches: k_nh_element (element matrices and residuals,
        # the tangent on the matrix cores, E to HBM) and k_gather_emat (node
        # gathers of E into the CSR).  The HBM roofline takes the compulsory
        # bytes over both kernels' time (traffic over the same two kernels);
        # the compute side is the element kernel's matrix-core rate.
        t2 = head["ms_geometry_kernel"] + head["ms_assemble_kernel"]
        r = out["roofline"]
        r.update(kernel="k_nh_element + k_gather_emat", kernel_ms=t2,
                 kernel_ms_min=None, achieved=compulsory / (t2 * 1e-3) / 1e9,
                 frac=compulsory / (t2 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 kernel_ms_stat="sum of the two kernels' medians over the timed steps (HIP events on the "
                                "launch stream)",
                 element_kernel_ms=head["ms_geometry_kernel"], gather_kernel_ms=head["ms_assemble_kernel"])
        r["element_scatter_model"]["note"] = ("SURVEY-style element-scatter count (a read, a write and a "
                                              "position per element-matrix entry, 16 B per residual entry); "
                                              "not used for frac")
        r["element_scatter_model"]["bytes_per_element"] = b_patch + 16 * n_loc
        r["element_scatter_model"]["gb_per_assembly"] = (b_patch + 16 * n_loc) * n_cv / 1e9
        # matrix-core products per element (both forms): 3 K3 node tiles x 21 and
        # 21 upper (component pair, node tile pair) tiles x 14 products of
        # 16x16x4 (2048 FLOP each), zero pads included; the spatial form's useful
        # FLOP: 48 per upper node-pair block and point
        n_mfma = 3 * 21 + 21 * 14
        useful = 48 * (27 * 28 // 2) * 27
        mf = {}
        for fname, instr, peak in (("fp64_element_batched", "v_mfma_f64_16x16x4_f64", FP64_PEAK_TFLOPS),
                                   ("fp32_element_batched", "v_mfma_f32_16x16x4_f32", FP32_PEAK_TFLOPS)):
            f = forms.get(fname)
            if not f:
                continue
            t_el = f["ms_geometry_kernel"] * 1e-3
            issued = n_mfma * 2048 * n_cv / t_el / 1e12
            mf[fname] = {"kernel": "k_nh_element", "instruction": instr, "mfma_per_element": n_mfma,
                         "issued_tflops": issued, "peak_tflops": peak, "frac": issued / peak,
                         "useful_flop_per_element": useful,
                         "useful_tflops": useful * n_cv / t_el / 1e12,
                         "kernel_ms": f["ms_geometry_kernel"],
                         "rel_frobenius_err_vs_fp64": f["rel_frobenius_err_vs_fp64"]}
        out["mfma"] = mf
        flop_qp = 2 * 27 * (n_loc * n_loc * 9 + n_loc * 81)
        out["per_qp_flop_model"] = {
            "flop_per_element": flop_qp,
            "equivalent_tflops": flop_qp * n_cv / (head["ms_geometry_kernel"] * 1e-3) / 1e12,
            "note": "per-point B^T D B count with d = 9 (2 nq (n_loc^2 d + n_loc d^2)); the spatial form needs "
                    "about a seventh of it, so this is the rate a per-point contraction would need, not an "
                    "issued rate"}
        out["dtype"] = "f64"
    return out, (pts, cvx, lam, mu, fem, q, im, term, state)


# ------------------------------------------------------------------ PMC leg
def _rocprof():
    for c in (shutil.which("rocprofv3"), "/opt/rocm/bin/rocprofv3"):
        if c and os.path.exists(c):
            return c
    return None


def _pmc_pass(counter, regex, child_args, timeout=180):
    """One rocprofv3 counter pass over a child run of this script; returns the
    per-dispatch counter values (KiB) of the kernels matching regex."""
    prof = _rocprof()
    if prof is None:
        raise RuntimeError("rocprofv3 not found")
    out = tempfile.mkdtemp(prefix="gfx_pmc_", dir="/tmp")
    try:
        cmd = [prof, "--pmc", counter, "--kernel-include-regex", regex, "-d", out, "-o", "pmc",
               "-f", "csv", "--", sys.executable, "-u", os.path.join(ROOT, "bench.py")] + child_args
        env = dict(os.environ, TMPDIR="/tmp")
        log("pmc pass:", counter, " ".join(child_args))
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           timeout=timeout)
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise RuntimeError("rocprofv3 (rc %d) wrote no counter file: %s" %
                               (r.returncode, r.stderr.decode(errors="replace")[-400:]))
        import csv
        vals = []
        pat = re.compile(regex)
        for f in files:
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name", counter) == counter and pat.search(row.get("Kernel_Name", "")):
                        vals.append(float(row["Counter_Value"]))
        return vals
    finally:
        shutil.rmtree(out, ignore_errors=True)


def measure_traffic(args):
    """HBM bytes per assembly of the dominant kernel(s): FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) of the listed kernels, per assembly as the
    difference between a child run of 3 assemblies and one of 1, halved (so
    launches of the same kernels during the pattern build cancel), each
    scaled by what the counter reads for a plain 1 GiB device copy."""
    regex = PMC_KERNELS[args.config]
    form = args.form if args.form != "all" else {"c3": "fp64_element_batched"}.get(args.config, "fp64_factorized")

    def child(steps):
        return ["--pmc-child", "--config", args.config, "--form", form, "--steps", str(steps), "--warmup", "0",
                "--no-cpu"]
    copy_bytes = float(8 << 27)
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = _pmc_pass(counter, "copy|elementwise", ["--copy-calibration"])
        if not cal:
            raise RuntimeError("no copy kernel seen in the calibration pass")
        factor = copy_bytes / (statistics.median(cal) * 1024.0)
        v1 = _pmc_pass(counter, regex, child(1))
        v3 = _pmc_pass(counter, regex, child(3))
        if not v1 or not v3:
            raise RuntimeError("no %s dispatch seen" % regex)
        res[counter] = {"kib_per_assembly": (sum(v3) - sum(v1)) / 2.0,
                        "dispatches_per_assembly": (len(v3) - len(v1)) / 2.0,
                        "copy_kib": statistics.median(cal), "factor": factor}
    traffic = sum(v["kib_per_assembly"] * 1024.0 * v["factor"] for v in res.values())
    return traffic, res


# -------------------------------------------------------------- CPU baseline
def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(share, int(omp)) if omp and omp.isdigit() else share
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "affinity_cpus": share,
            "omp_num_threads_env": omp, "threads_used": threads}


def _median_runs(fn, runs=5):
    fn()  # warm-up
    t = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return statistics.median(t), t


def cpu_baseline(inputs, config, info):
    """The oracle (a C port of GetFEM's per-element/per-point FP64 algorithm
    with its threshold and sorted-column insertion, OpenMP over elements with
    per-thread matrices summed as model::assembly does) on bounded samples of
    the same workload: 1 warm-up + median of 5 on the CPU share, 1 warm-up +
    median of 3 on one core."""
    import oracle
    pts, cvx, lam, mu, fem, q, im, term, state = inputs
    k = int(fem[-2])
    if term == "neohookean":
        dofs, ndof = oracle.enumerate_dof(oracle.GT_QK31, pts, cvx, oracle.FEM_QK, k, q)
        prob = oracle.Problem(oracle.GT_QK31, pts, cvx, oracle.FEM_QK, k, q, dofs, ndof, im,
                              oracle.TERM_NEOHOOKEAN, scal_a=lam, scal_b=mu, U=state)
    else:
        dofs, ndof = oracle.enumerate_dof(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, q)
        if term == "laplace":
            prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, q, dofs, ndof, im,
                                  oracle.TERM_LAPLACE, scal_a=1.0)
        else:
            prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, q, dofs, ndof, im,
                                  oracle.TERM_ISO_ELAST, coef_a=lam, coef_b=mu)
    n_cv = cvx.shape[0]
    th = info["threads_used"]
    # samples sized for about 2.5 s per run on 16 threads and 1 s on one core
    # (measured rates: P2 45k / 100k elems/s, P3 5.2k / 19k, Q2 1.3k / 7.4k)
    many, one = {"c1": (48000, 48000), "c2": (250000, 45000), "c5": (48000, 5000),
                 "c3": (18000, 1300)}[config]
    many = min(n_cv, max(1000, int(many * th / 16)))
    one = min(n_cv, one)
    tm, runs_m = _median_runs(lambda: prob.assemble_timed(0, many, th), 5)
    t1, runs_1 = _median_runs(lambda: prob.assemble_timed(0, one, 1), 3)
    log("cpu baseline: %d threads %d elements median %.2f s; 1 thread %d elements median %.2f s" %
        (th, many, tm, one, t1))
    return {
        "value": many / tm, "unit": "elems/s", "cores": th, "kind": "port",
        "sample": "first %d of %d elements per run, 1 warm-up + median of 5 runs (%s s) on %d threads "
                  "(the host's CPU share: %d CPUs in the affinity mask, OMP_NUM_THREADS=%s, os.cpu_count()=%s, "
                  "%s); OpenMP element partition with per-thread matrices summed"
                  % (many, n_cv, ", ".join("%.2f" % x for x in runs_m), th, info["affinity_cpus"],
                     info["omp_num_threads_env"], info["os_cpu_count"], info["cpu_model"]),
        "value_1core": one / t1,
        "sample_1core": "first %d elements, 1 warm-up + median of 3 runs (%s s)"
                        % (one, ", ".join("%.2f" % x for x in runs_1)),
        "cpu": info,
        "note": "the port is a faithful restatement, not GetFEM itself: GetFEM measured on an 8-vCPU Xeon "
                "(1 thread) runs C2 at 8,090 elems/s, C3 at 667, C5 at 1,105, so this port is about 5.7x "
                "(C2) faster per core than the reference's own CPU path",
    }


def run_c4(args):
    """SIMP compliance J = u^T K(rho) u on P1 tets (32^3x6) with its three
    gradients (dJ/drho, dJ/dX, dJ/du) in one device call per step."""
    import torch

    import getfem_amd as ga
    from getfem_amd import meshes
    from getfem_amd.torch_ops import Compliance

    n = args.n or 32
    nu = 0.3
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1 / (2 * (1 + nu))
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
    op.check()
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
                               "call (element kernel + one launch for the two node gathers and the J sum), "
                               "replayed from a HIP graph" % (n, n_cv),
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


def run_c2solve(args):
    """The C2 system clamped on the face x = 0 (Dirichlet by simplification on
    the device CSR), right-hand side K applied to a smooth displacement (so
    the solution is known), solved by the device CG with GetFEM's default
    preconditioner for coercive problems (ILDLT, rebuilt at every solve) and
    with block Jacobi: iterations, times and the error against the known
    solution."""
    import torch

    import getfem_amd as ga
    from getfem_amd import _lib, meshes

    n = args.n or 64
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    lam, mu = meshes.lognormal_lame(cvx.shape[0])
    torch.cuda.init()
    dev = torch.device("cuda:0")
    ctx = ga.Context(0)
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, [lam, mu])
    nnz = plan.build_pattern()
    plan.assemble()
    nodes = mf.basic_dof_nodes()  # one row per node; dof = 3 node + component
    ndof = mf.nbdof()
    on_face = np.flatnonzero(np.abs(nodes[:, 0]) < 1e-12)
    fixed = (3 * on_face[:, None] + np.arange(3)[None, :]).ravel().astype(np.int32)
    x_true = meshes.sine_modes()(nodes, 0.05).ravel()
    x_true[fixed] = 0.0
    dxt = torch.tensor(x_true, device=dev)
    db = torch.empty_like(dxt)
    plan.dof_constraints(fixed, np.zeros(fixed.size),
                         flags=_lib.BUILD_MATRIX | _lib.MODEL_LINEAR | _lib.MODEL_SYMMETRIC)
    plan.spmv(dxt.data_ptr(), db.data_ptr())
    torch.cuda.synchronize()
    out = {"metric": METRIC, "unit": "iterations", "n_gpus": 1, "higher_is_better": False,
           "dtype": "f64", "data": "synthetic (seeded jittered mesh, lognormal lambda/mu)",
           "config": {"workload": "P2-tet isotropic elasticity %d^3x6 = %d tets, %d dofs, %d stored entries, "
                                  "clamped on x = 0 (%d dofs), b = K x_true, x_true = sum of 8 sine modes; "
                                  "device CG to |r| <= %g |b| from x = 0" % (n, cvx.shape[0], ndof, nnz,
                                                                             fixed.size, args.rtol),
                      "name": "c2solve"},
           "solves": {}}
    t0 = time.perf_counter()
    nf, nb = plan.ildlt_factor()
    torch.cuda.synchronize()
    out["ildlt_levels"] = {"forward": nf, "backward": nb, "units": ndof // 3,
                           "first_factorization_s": time.perf_counter() - t0}
    for name, pc in (("ildlt", ga.PRECOND_ILDLT), ("block_jacobi", ga.PRECOND_BLOCK_JACOBI)):
        dx = torch.zeros_like(dxt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, res = plan.solve_pcg(db.data_ptr(), dx.data_ptr(), pc, rtol=args.rtol, maxit=args.maxit)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        err = float(torch.linalg.vector_norm(dx - dxt) / torch.linalg.vector_norm(dxt))
        out["solves"][name] = {"iterations": it, "rel_residual": res, "seconds": t,
                               "ms_per_iteration": 1e3 * t / max(it, 1), "rel_error_vs_x_true": err}
        log(name, json.dumps(out["solves"][name]))
    out["value"] = out["solves"]["ildlt"]["iterations"]
    return out


def cpu_baseline_c4(inputs):
    """The oracle's compliance with its analytic gradients over the whole
    mesh: 1 warm-up + median of 5 (one thread)."""
    import oracle
    pts, cvx, dofs, ndof, lam, mu, rho, u = inputs
    prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, 1, 3, dofs, ndof, "IM_TETRAHEDRON(1)",
                          oracle.TERM_ISO_ELAST, scal_a=lam, scal_b=mu, U=u)
    t, runs = _median_runs(lambda: prob.compliance(rho, 1e-3, 3.0, grads=True), 5)
    return {"value": cvx.shape[0] / t, "unit": "elems/s", "cores": 1, "kind": "port",
            "sample": "the whole %d-element mesh (J, dJ/drho, dJ/dX), 1 warm-up + median of 5 runs (%s s)"
                      % (cvx.shape[0], ", ".join("%.3f" % x for x in runs)), "cpu": cpu_info()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c4", "c2solve"])
    ap.add_argument("--rtol", type=float, default=1e-8, help="c2solve: CG tolerance")
    ap.add_argument("--maxit", type=int, default=20000, help="c2solve: CG iteration cap")
    ap.add_argument("--n", type=int, default=0, help="mesh subdivisions (default: the config's)")
    ap.add_argument("--form", default="all")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 HBM traffic passes")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
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
    if args.config == "c2solve":
        print(json.dumps(run_c2solve(args)), flush=True)
        return
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
    if args.pmc_child:
        return
    n0 = CONFIGS[args.config][0]
    if not args.no_pmc and args.config in PMC_KERNELS and (args.n or n0) == n0:
        try:
            traffic, detail = measure_traffic(args)
            r = out["roofline"]
            r["traffic"] = traffic
            t_ms = r["kernel_ms"]
            r["traffic_gbs"] = traffic / (t_ms * 1e-3) / 1e9
            r["traffic_frac"] = r["traffic_gbs"] / HBM_PEAK_GBS
            r["traffic_pmc"] = detail
            r["traffic_kernels"] = PMC_KERNELS[args.config]
            r["traffic_note"] = ("measured in this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over "
                                 "child runs of the same workload and form (3 and 1 assemblies, difference "
                                 "halved), summed over one assembly's launches of the listed kernels, each "
                                 "scaled by the factor a plain 1 GiB device copy gives")
        except Exception as exc:  # report, never hide
            out["roofline"]["traffic_error"] = repr(exc)
    if not args.no_cpu:
        try:
            import oracle
            if not os.path.exists(oracle.LIB_PATH):
                oracle.build()
            out["cpu_baseline"] = cpu_baseline(inputs, args.config, cpu_info())
        except Exception as exc:  # report, never hide
            out["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
