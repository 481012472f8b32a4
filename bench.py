"""Benchmark: solve time of the dense mollified-BEM system on one MI355X.

Workload (default): the box-with-dipping-fault model refined to N = 20,000
box triangles (n = 60,000 unknowns, the FP64 matrix is 28.8 GB) plus a
676-triangle fault with uniform unit reverse slip, eps = 0.5 h, mu = 30 GPa,
nu = 0.25 (synthetic geometry generated here; no data files).

One step = the whole FP64 GPU solve from the mesh arrays resident in HBM:
dense assembly (calibrated jump, fused row sums), the block-Jacobi
preconditioner build (patch super-blocks forced to the block-Jacobi rung,
9,000-DOF chunks, explicit chunk inverses), and FGMRES (restart 200) to the
relative error target against an FP64 reference solution computed once,
untimed, at rtol 1e-13.

Prints one JSON line (see the keys at the end of ``main``).  ``--gpus`` above
1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solve time to 1e-8 rel err (N=20k tris): mixed+learned vs plain FP64 GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def _measured_traffic_ratio():
    """HBM bytes / algorithmic bytes of the matvec kernel from the committed
    rocprofv3 PMC summary (FETCH_SIZE / WRITE_SIZE scaled on a plain copy),
    or None when no such summary is present."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "matvec_traffic_n*.txt"))):
        m = re.search(r"traffic_over_algorithmic = ([0-9.eE+-]+)", open(path).read())
        if m:
            return float(m.group(1)), os.path.basename(path)
    return None, None


def _events_time(fn, reps):
    import torch
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def cpu_baseline(model, eps, system, n_iters, n_rows=48, gemv_rows=3000):
    """The oracle (C restatement, OpenMP on every host core this process may
    use) timed on a bounded sample of the same workload, extrapolated to the
    whole job the reference's dense path runs: T blocks for every (q, p) pair
    (the reference row-sums all of them), U blocks for every row against the
    prescribed-displacement sources, the preconditioner's chunk LUs, and
    ``n_iters`` + 2 FGMRES matvecs plus one preconditioner apply each."""
    import scipy.linalg as sla
    from oracle import tri_kernels as ok
    from oracle.dense import _arrays, _t_coeffs, _u_coeffs
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    region = model.regions[0]
    mat = region.material
    pts_all = np.concatenate([p.collocation_points() for p in region.patches])
    rng = np.random.default_rng(0)
    rows = pts_all[rng.choice(pts_all.shape[0], n_rows, replace=False)]
    t_src = [(_arrays(p), eps[p.name]) for p in region.patches]
    u_src = [(_arrays(p), eps[p.name]) for p in region.patches if p.bc.name == "PRESCRIBED_DISPLACEMENT"]
    c = _t_coeffs(mat.mu, mat.lam)
    g = _u_coeffs(mat.mu, mat.lam)
    t0 = time.perf_counter()
    nT = 0
    for (tv, nrm), e in t_src:
        ok.t_matrix_direct(rows, tv, nrm, np.full(tv.shape[0], e), *c)
        nT += rows.shape[0] * tv.shape[0]
    tT = (time.perf_counter() - t0) / nT
    t0 = time.perf_counter()
    nU = 0
    for (tv, nrm), e in u_src:
        ok.u_matrix_direct(rows, tv, np.full(tv.shape[0], e), *g)
        nU += rows.shape[0] * tv.shape[0]
    tU = (time.perf_counter() - t0) / nU
    N = pts_all.shape[0]
    n = 3 * N
    n_dir = sum(p.n_triangles for p in region.patches if p.bc.name == "PRESCRIBED_DISPLACEMENT")
    assembly_s = tT * N * N + tU * N * n_dir
    # matvec: numpy GEMV (threaded BLAS) on a row sample of an n-column matrix
    Ablk = rng.standard_normal((gemv_rows, n))
    xv = rng.standard_normal(n)
    t0 = time.perf_counter()
    for _ in range(3):
        Ablk @ xv
    gemv_s = (time.perf_counter() - t0) / 3 * (n / gemv_rows)
    # preconditioner: LU of one chunk-sized block, scaled by m^3 to every chunk
    m = 1500
    Bm = rng.standard_normal((m, m)) + m * np.eye(m)
    t0 = time.perf_counter()
    lu = sla.lu_factor(Bm)
    lu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    sla.lu_solve(lu, xv[:m])
    trs_s = time.perf_counter() - t0
    chunk = 7500 if n >= 60000 else min(9000, n // 4)
    n_chunks = max(1, n // chunk)
    build_s = n_chunks * lu_s * (chunk / m) ** 3
    apply_s = n_chunks * trs_s * (chunk / m) ** 2 + 0.3 * gemv_s
    total = assembly_s + build_s + (n_iters + 2) * gemv_s + n_iters * apply_s
    sample = (f"oracle C pair kernels on {n_rows} sampled rows x all {N} sources (T) and "
              f"{n_dir} prescribed sources (U) = {nT + nU:,} pairs, extrapolated to the "
              f"reference's {N * N + N * n_dir:,} pairs; numpy GEMV on {gemv_rows} of {n} rows "
              f"x {n_iters + 2} matvecs; scipy LU of a {m}^2 block scaled by (m/{m})^3 to "
              f"{n_chunks} chunks of {chunk} DOFs")
    return {"value": total, "unit": "s", "cores": cores, "kind": "port", "sample": sample,
            "parts_s": {"assembly": assembly_s, "precond_build": build_s,
                        "fgmres": (n_iters + 2) * gemv_s + n_iters * apply_s}}


def kernel_probe(n: int, reps: int):
    """Traffic probe for the PMC counters: ``reps`` launches of the FGMRES
    matvec kernel on an n x n FP64 matrix, then ``reps`` plain device copies
    of the same byte count (the counter calibration: a copy of B bytes reads
    B and writes B).  Prints the per-launch algorithmic bytes."""
    import torch
    from mbem_amd.la.solver import gemv
    dev = torch.device("cuda", 0)
    A = torch.randn((n, n), dtype=torch.float64, device=dev)
    x = torch.randn(n, dtype=torch.float64, device=dev)
    y = torch.empty(n, dtype=torch.float64, device=dev)
    for _ in range(reps):
        gemv(A, x, y)
    torch.cuda.synchronize()
    B = torch.empty_like(A)
    for _ in range(reps):
        B.copy_(A)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "gemv+copy", "n": n, "gemv_bytes": n * n * 8 + 2 * n * 8,
                      "copy_bytes_each_way": n * n * 8, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-cells", type=int, default=50)
    ap.add_argument("--fault-cells", type=int, nargs=2, default=(13, 26))
    ap.add_argument("--target-err", type=float, default=1e-8)
    ap.add_argument("--rtol", type=float, default=2e-11)
    ap.add_argument("--bj-chunk", type=int, default=3000)
    ap.add_argument("--grouping", choices=("patch", "region"), default="region")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--method", choices=("plain", "mixed", "both"), default="both")
    ap.add_argument("--storage", choices=("bf16", "fp32"), default="bf16")
    ap.add_argument("--inner-rtol", type=float, default=5e-2)
    ap.add_argument("--bj-chunk-mixed", type=int, default=3000)
    ap.add_argument("--grouping-mixed", choices=("patch", "region"), default="region")
    ap.add_argument("--probe-traffic", type=int, default=0, metavar="N",
                    help="only run the matvec / copy traffic probe on an N x N matrix")
    args = ap.parse_args()
    if args.probe_traffic:
        kernel_probe(args.probe_traffic, 3)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(the dense path's sharding is described, not run)", file=sys.stderr)
        sys.exit(2)
    import torch
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.cases import box_fault_model
    from mbem_amd.la.solver import gemv
    from mbem_amd.model import generate_system

    t_setup = time.perf_counter()
    model, eps, info = box_fault_model(args.n_cells, fault_cells=tuple(args.fault_cells))
    system = generate_system(model)
    n = system.layout.n_unknowns
    print(f"[bench] workload n_cells={args.n_cells} N={info['box_triangles']} n={n} "
          f"fault={info['fault_triangles']}", flush=True)

    def step():
        asm = AssembledDense(system, eps, "direct", jump="calibrated")
        x = asm.solve(rtol=args.rtol, precond_max_dense=0, precond_bj_chunk=args.bj_chunk,
                      return_device=True, precond_grouping=args.grouping)
        return asm, x

    asm = None
    for w in range(args.warmup):
        asm = None
        torch.cuda.empty_cache()
        asm, _ = step()
        print(f"[bench] warmup {w}: assembly {asm.assembly.seconds:.3f} s, {asm.report}",
              flush=True)
    if asm is None:
        asm = AssembledDense(system, eps, "direct", jump="calibrated")
    asm.release()
    x_ref = asm.solve(rtol=1e-13, precond_max_dense=0, precond_bj_chunk=args.bj_chunk,
                      return_device=True).clone()
    ref_report = asm.report
    print(f"[bench] reference solve: {ref_report}", flush=True)
    # roofline of the FGMRES matvec, timed live with events on the stream it runs on
    A = asm.A_dev
    y = torch.empty(n, dtype=torch.float64, device=A.device)
    gemv_ms = _events_time(lambda: gemv(A, x_ref, y), 10)
    gemv_bytes = n * n * 8 + 2 * n * 8
    asm = None
    torch.cuda.empty_cache()
    print(f"[bench] gemv {gemv_ms:.3f} ms = {gemv_bytes / gemv_ms / 1e6:.0f} GB/s setup "
          f"{time.perf_counter() - t_setup:.1f} s", flush=True)

    def timed(step_fn, label):
        times, errs, iters = [], [], []
        asm_k = None
        torch.cuda.synchronize()
        for k in range(args.steps):
            asm_k = None
            torch.cuda.empty_cache()
            ts = time.perf_counter()
            asm_k, x, rep = step_fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - ts)
            errs.append(float(torch.linalg.vector_norm(x - x_ref)
                              / torch.linalg.vector_norm(x_ref)))
            iters.append(rep.iterations)
            ph = ", ".join(f"{k_}={v:.3f}" for k_, v in asm_k.timings.items())
            print(f"[bench] {label} step {k}: {times[-1]:.3f} s ({ph}; {rep}; "
                  f"rel err {errs[-1]:.2e})", flush=True)
        return times, errs, iters, asm_k, rep

    def step_plain():
        a, x = step()
        return a, x, a.report

    def step_mixed():
        from mbem_amd.la.mixed import solve_mixed
        a = AssembledDense(system, eps, "direct", jump="calibrated")
        t0 = time.perf_counter()
        x, rep = solve_mixed(a, rtol=args.rtol, storage=args.storage,
                             inner_rtol=args.inner_rtol, precond_bj_chunk=args.bj_chunk_mixed,
                             precond_grouping=args.grouping_mixed)
        a.timings.update(mixed_operator_s=rep.operator_build_s,
                         precond_s=rep.precond_build_s, fgmres_s=rep.solve_s)
        return a, x, rep

    results = {}
    if args.method in ("mixed", "both"):
        step_mixed()          # warm the mixed kernels and rocSOLVER's FP32 path
        torch.cuda.synchronize()
    for label, fn in (("plain", step_plain), ("mixed", step_mixed)):
        if args.method in (label, "both"):
            results[label] = timed(fn, label)
    main_label = "mixed" if "mixed" in results else "plain"
    times, errs, iters, asm, rep_main = results[main_label]
    total = float(np.sum(times))
    ms_per_step = 1e3 * total / args.steps
    plain_s = float(np.mean(results["plain"][0])) if "plain" in results else None
    mixed_s = float(np.mean(results["mixed"][0])) if "mixed" in results else None
    pre = results["plain"][4].precond_summary if "plain" in results else {}
    cpu = None
    if not args.no_cpu_baseline:
        it_plain = results["plain"][2] if "plain" in results else [60]
        cpu = cpu_baseline(model, eps, system, int(np.median(it_plain)))
        print(f"[bench] cpu baseline {cpu['value']:.1f} s on {cpu['cores']} cores", flush=True)
    ratio, src_name = _measured_traffic_ratio()
    line = {
        "metric": METRIC,
        "value": ms_per_step / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic box-with-dipping-fault mesh generated in bench.py",
        "config": {
            "workload": f"box 40x40x20 km, N={info['box_triangles']} P0 triangles "
                        f"(n={n}), dipping fault {info['fault_triangles']} triangles, "
                        f"eps=0.5h, mu=30 GPa nu=0.25",
            "method": ("mixed: BF16 far field + FP64 near field on an FP64 outer FGMRES, "
                       "FP32 block-Jacobi" if main_label == "mixed" else
                       "plain FP64: GPU assembly + block-Jacobi FGMRES"),
            "storage": args.storage if main_label == "mixed" else "fp64",
            "rtol": args.rtol, "rel_err_vs_fp64_reference": max(errs),
            "target_rel_err": args.target_err, "iterations": iters,
            "plain_fp64_solve_s": plain_s, "mixed_solve_s": mixed_s,
            "speedup_vs_plain_fp64": (plain_s / mixed_s if plain_s and mixed_s else None),
            "plain_rel_err": (max(results["plain"][1]) if "plain" in results else None),
            "precond_build_s": pre.get("build_s"), "precond_chunks": pre.get("chunks"),
            "plain_precond": f"block-Jacobi FP64, {args.grouping} grouping, chunk {args.bj_chunk}",
            "mixed_precond": (f"inner FGMRES to {args.inner_rtol:g} on the mixed operator, "
                              f"block-Jacobi FP32, {args.grouping_mixed} grouping, chunk "
                              f"{args.bj_chunk_mixed}"),
            "learned_preconditioner": "not built yet",
        },
        "roofline": {"bound": "hbm", "achieved": gemv_bytes / gemv_ms / 1e6,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gemv_bytes / gemv_ms / 1e6 / HBM_PEAK_GBS,
                     "traffic": (None if ratio is None else ratio * gemv_bytes),
                     "traffic_source": src_name,
                     "kernel": "k_gemv (FGMRES matvec, n x n FP64)",
                     "algorithmic_bytes": gemv_bytes, "avg_ms": gemv_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
