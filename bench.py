"""Benchmark: solve time of the dense mollified-BEM system on one MI355X.

Workload (default): the box-with-dipping-fault model refined to N = 20,000
box triangles (n = 60,000 unknowns, the FP64 matrix is 28.8 GB) plus a
676-triangle fault with uniform unit reverse slip, eps = 0.5 h, mu = 30 GPa,
nu = 0.25 (synthetic geometry generated here; no data files).

One step = a whole GPU solve from the mesh arrays: dense assembly
(calibrated jump, fused row sums), the preconditioner build (block-Jacobi
over cluster chunks, explicit chunk inverses) and the Krylov solve to the
relative-error target against an FP64 reference solution computed once,
untimed, at rtol 1e-13.  Two methods are stepped: plain FP64 (block-Jacobi
FGMRES) and mixed (int8 far field + FP64 near field inner FGMRES with FP32
block-Jacobi, a rigid coarse space and a recycled Krylov space, under an
FP64 outer FGMRES whose operator is, by default, the int8 codes plus their
24-bit rounding remainders: every far entry within 2e-10 of its segment's
largest, from 4 bytes instead of 8); ``value`` is the mixed step, the plain
step and the ratio are in ``config``.
The learned preconditioner's A/B on the same system is reported beside it.

``--workload laplace`` (batched Laplace-state solves) and ``--workload
readout`` (stress-readout throughput) print their own JSON lines.
``--gpus`` above 1 exits: bench.py measures one GPU.
"""
from __future__ import annotations

import argparse
import gc
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
FP32_PEAK_TFLOPS = 157.3                # dense FP32 matrix-core peak (sgetrf roofline)


def _measured_traffic_ratio(kernel: str = "k_gemv"):
    """HBM bytes / algorithmic bytes of a matvec kernel from the committed
    rocprofv3 PMC summary (FETCH_SIZE / WRITE_SIZE scaled on a plain copy),
    or None when no such summary is present."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "matvec_traffic_n*.txt"))):
        m = re.search(r"^" + re.escape(kernel) + r" traffic_over_algorithmic = ([0-9.eE+-]+)",
                      open(path).read(), re.M)
        if m:
            return float(m.group(1)), os.path.basename(path)
    return None, None


def _executed_flops():
    """{kernel: executed FP64 flops per pair} from the committed PMC summary."""
    import re
    path = os.path.join(ROOT, "profiles", "fp64_executed_flops.md")
    if not os.path.exists(path):
        return {}
    return {m.group(1): float(m.group(2)) for m in
            re.finditer(r"executed_flops_per_pair (\w+) = ([0-9.]+)", open(path).read())}


def i8_bytes(n: int) -> int:
    """Algorithmic bytes of one int8 far-field matvec (n x n): the codes, one
    FP32 scale per row and 64 columns, x in FP32, y in FP64."""
    return n * n + 4 * n * (-(-n // 64)) + 4 * n + 8 * n


def i8e_bytes(n: int) -> int:
    """Algorithmic bytes of one split far-field matvec (n x n): the int8
    codes, the 24-bit remainders, one FP32 scale per row and 64 columns, x
    and y in FP64."""
    return n * n + 3 * n * n + 4 * n * (-(-n // 64)) + 8 * n + 8 * n


def _release(full: bool = False):
    """Free what the last step left: collect garbage (a dropped object still
    in a reference cycle keeps its device memory until then); with ``full``
    also return the allocator's cached blocks to the device.  Between steps
    of one method the blocks stay cached, so each step reuses the last one's
    (as any caller solving repeatedly does) instead of mapping 28.8 GB
    afresh; between methods and phases the cache is emptied."""
    import torch
    gc.collect()
    if full:
        torch.cuda.empty_cache()


def quantity_errors(model, system, eps, x, x_ref):
    """Relative errors of a solution ``x`` against the FP64 reference
    ``x_ref`` (both device vectors over the system's unknowns), per quantity:
    the u-slots (displacement), the t-slots (traction), and the subtracted
    stress read on the 50 x 50 grid over the top face (z = 0) and 4 eps of
    the top below it.  Each is max |difference| / max |reference|."""
    import torch
    from mbem_amd.cases import surface_grid
    from mbem_amd.evaluate import evaluate_stress
    layout = system.layout
    xh, rh = x.cpu().numpy(), x_ref.cpu().numpy()
    out = {}
    for kind in ("u", "t"):
        idx = np.concatenate([np.arange(s_.offset, s_.stop) for s_ in layout.slots
                              if s_.kind == kind] or [np.zeros(0, int)])
        if idx.size:
            out[kind] = float(np.abs(xh[idx] - rh[idx]).max() / np.abs(rh[idx]).max())
    region = model.regions[0]
    sol = {s_.name: xh[s_.offset:s_.stop].reshape(-1, 3) for s_ in layout.slots}
    ref = {s_.name: rh[s_.offset:s_.stop].reshape(-1, 3) for s_ in layout.slots}
    top_eps = float(eps["top"]) if isinstance(eps, dict) else float(eps)
    for tag, z in (("stress_z0", 0.0), ("stress_4eps", -4.0 * top_eps)):
        pts = surface_grid(50, half=20.0, z=z)
        sa = evaluate_stress(model, region, sol, pts, eps, warn_near=False)
        sb = evaluate_stress(model, region, ref, pts, eps, warn_near=False)
        out[tag] = float(np.abs(sa - sb).max() / np.abs(sb).max())
    torch.cuda.synchronize()
    return out


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
    matvec kernel on an n x n FP64 matrix, ``reps`` launches of the mixed
    method's BF16 far-field matvec on its BF16 copy, of the int8 one on its
    int8 copy and of the split outer matvec (int8 codes + 24-bit remainders,
    FP64 x), then ``reps`` plain device copies of the FP64 matrix (the
    counter calibration: a copy of B bytes reads B and writes B).  Prints the
    per-launch algorithmic bytes."""
    import torch
    from mbem_amd._lib import call, load, ptr, stream
    from mbem_amd.la.solver import gemv
    dev = torch.device("cuda", 0)
    A = torch.randn((n, n), dtype=torch.float64, device=dev)
    x = torch.randn(n, dtype=torch.float64, device=dev)
    y = torch.empty(n, dtype=torch.float64, device=dev)
    for _ in range(reps):
        gemv(A, x, y)
    torch.cuda.synchronize()
    ldf = (n + 7) // 8 * 8
    F = torch.empty((n, ldf), dtype=torch.int16, device=dev)
    call("mbem_cast_lo", ptr(A), n, n, n, ptr(F), ldf, 1, stream())
    xf = x.float()
    for _ in range(reps):
        call("mbem_gemv_lo", ptr(F), ldf, n, n, ptr(xf), ptr(y), 1.0, 0.0, 1, stream())
    torch.cuda.synchronize()
    del F
    ld8 = (n + 15) // 16 * 16
    nseg = int(load().mbem_i8_segments(n))
    F8 = torch.empty((n, ld8), dtype=torch.uint8, device=dev)
    S8 = torch.empty((n, nseg), dtype=torch.float32, device=dev)
    ldr = int(load().mbem_i8e_resid_bytes(n))
    E = torch.empty((n, ldr), dtype=torch.uint8, device=dev)
    call("mbem_cast_i8e", ptr(A), n, n, n, ptr(F8), ld8, ptr(S8), nseg, ptr(E), ldr, stream())
    for _ in range(reps):
        call("mbem_gemv_i8", ptr(F8), ld8, ptr(S8), nseg, n, n, ptr(xf), ptr(y), 1.0, 0.0,
             stream())
    torch.cuda.synchronize()
    for _ in range(reps):
        call("mbem_gemv_i8e", ptr(F8), ld8, ptr(E), ldr, ptr(S8), nseg, n, n, ptr(x), ptr(y),
             1.0, 0.0, stream())
    torch.cuda.synchronize()
    del F8, S8, E
    B = torch.empty_like(A)
    for _ in range(reps):
        B.copy_(A)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "gemv+gemv_lo+gemv_i8+gemv_i8e+copy", "n": n,
                      "gemv_bytes": n * n * 8 + 2 * n * 8,
                      "gemv_lo_bytes": n * n * 2 + n * 4 + n * 8,
                      "gemv_i8_bytes": i8_bytes(n), "gemv_i8e_bytes": i8e_bytes(n),
                      "copy_bytes_each_way": n * n * 8, "reps": reps}), flush=True)


def bench_laplace(args):
    """Batched Laplace-state solves: one relaxing Maxwell region (eta = 1e19
    Pa s, mu = 30 GPa, tau = eta/mu), 32 s values from s_grid(0.01 tau,
    2e7 tau), 8 random unit slip fields (seed 5) on the fault, every
    (state, slip) pair solved in one lockstep batch (K = 256 columns) to
    rtol 1e-10.  Prints one JSON line of its own."""
    import torch
    from mbem_amd import laplace, rheology as rh
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.cases import box_fault_model
    from mbem_amd.kernels import basis as kb
    from mbem_amd.model import generate_system
    from mbem_amd.viscoelastic import LaplaceStates

    n_cells = args.laplace_cells
    model, eps, info = box_fault_model(n_cells, fault_cells=tuple(args.laplace_fault_cells))
    region = model.regions[0]
    h = float(np.mean([kb.element_sizes(p.mesh).mean() for p in region.patches]))
    eps = {p.name: 0.5 * h for p in region.patches}
    eps["fault"] = 0.5 * float(kb.element_sizes(region.faults[0].mesh).mean())
    system = generate_system(model)
    n = system.layout.n_unknowns
    tau = 1.0e19 / 3.0e10
    grid = laplace.s_grid(0.01 * tau, 2e7 * tau)
    states = rh.laplace_states({"crust": (30.0, 30.0)}, {"crust": rh.Maxwell(tau)}, grid)[1:]
    rng = np.random.default_rng(5)
    nf = region.faults[0].mesh.n_triangles
    slips = []
    for _ in range(args.laplace_rhs):
        v = rng.normal(size=(nf, 3))
        slips.append({"fault": v / np.linalg.norm(v, axis=1, keepdims=True)})
    print(f"[laplace] N={info['box_triangles']} n={n} fault={nf} states={len(states)} "
          f"rhs={len(slips)}", flush=True)
    times = []
    x = rep = ls = None
    for k in range(args.warmup + args.steps):
        ls = None
        _release()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ls = LaplaceStates(system, eps, {"crust": rh.Maxwell(tau)}, bj_chunk=args.bj_chunk)
        x, rep = ls.solve(states, slips, rtol=args.laplace_rtol, restart=args.laplace_restart,
                          n_anchors=args.laplace_anchors)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        tag = "warmup" if k < args.warmup else "step"
        print(f"[laplace] {tag} {k}: {dt:.3f} s ({rep})", flush=True)
        if k >= args.warmup:
            times.append(dt)
    # roofline of the batched operator GEMM, timed live with events on its stream
    from mbem_amd.la.batched import GOperator
    K = len(states) * len(slips)
    K += K & 1
    g = torch.full((K,), 1.5, dtype=torch.float64, device="cuda")
    op = GOperator(ls.A[0], ls.A[1], g, pair=(ls.AA, ls.L))
    op._prepare()
    op.cfg = rep.gemm_cfg["gemm_g"]
    from mbem_amd.la import autotune
    tuned = autotune.report(f"gemm_g n={n} K={K} lib=True")
    X = torch.randn(n, K, dtype=torch.float64, device="cuda")
    Y = torch.empty_like(X)
    gemm_ms = _events_time(lambda: op.apply(X, Y), 5)
    flops = op.flops
    # sequential comparison on sampled states: the per-state path of the
    # reference (re-assembly at the state's moduli, then a solve per slip)
    seq = []
    errs = []
    pick = sorted({0, len(states) // 2, len(states) - 1})[:args.laplace_seq_states]
    b0, b1 = ls.rhs_parts(slips)
    for i in pick:
        st = states[i]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        asm = AssembledDense(system, eps, _materials=dict(st.materials))
        mu, lam, gval = ls.state_moduli(st.s)
        M = asm.preconditioner(0, args.bj_chunk, grouping="region")
        from mbem_amd.la.solver import fgmres
        xs = []
        for r in range(len(slips)):
            b = (b0[:, r] + gval * b1[:, r]).contiguous()
            xr, srep = fgmres(asm, b, M=M, rtol=args.laplace_rtol, restart=200, maxiter=600)
            xs.append(xr)
        torch.cuda.synchronize()
        seq.append(time.perf_counter() - t0)
        for r in range(len(slips)):
            ref = xs[r].cpu().numpy()
            errs.append(float(np.linalg.norm(x[i, r] - ref) / np.linalg.norm(ref)))
        del asm, M
        _release(full=True)
        print(f"[laplace] sequential state {st.label}: {seq[-1]:.3f} s, "
              f"max rel diff vs batched {max(errs[-len(slips):]):.2e}", flush=True)
    seq_est = float(np.mean(seq)) * len(states)
    value = float(np.mean(times))
    line = {
        "metric": "batched Laplace-state solve time (32 s x 8 RHS, one Maxwell region)",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * value, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic box-with-dipping-fault mesh generated in bench.py",
        "config": {
            "workload": f"box 40x40x20 km, N={info['box_triangles']} P0 triangles (n={n}), "
                        f"fault {nf} triangles, Maxwell tau = 1e19 Pa s / 30 GPa, "
                        f"{len(states)} s values x {len(slips)} random slips",
            "rtol": args.laplace_rtol, "restart": args.laplace_restart,
            "anchors": args.laplace_anchors, "bj_chunk": args.bj_chunk,
            "assembly_s": rep.assembly_s, "precond_s": rep.precond_s, "solve_s": rep.solve_s,
            "lockstep_steps": rep.batched.steps, "cycles": rep.batched.cycles,
            "max_relres": float(rep.batched.relres.max()),
            "gemm_cfg": rep.gemm_cfg,
            "gemm_candidates_ms": tuned["ms"] if tuned else None,
            "gemm_library_id": op.LIB,
            "sequential_states_timed": len(pick),
            "sequential_per_state_s": [float(v) for v in seq],
            "sequential_estimate_s": seq_est,
            "speedup_vs_sequential": seq_est / value,
            "max_rel_diff_vs_sequential": max(errs) if errs else None,
        },
        "roofline": {"bound": "mfma", "achieved": flops / gemm_ms / 1e9,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": flops / gemm_ms / 1e9 / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": op.describe() + " (FP64 MFMA)",
                     "algorithmic_flops": flops, "avg_ms": gemm_ms},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_readout(args):
    """Stress-readout throughput: dd_stress_contract - eigenstress_contract,
    FP64, full sum, 10^6 points x 20,676 sources.  Prints one JSON line."""
    import torch
    from mbem_amd.kernels import device as dv
    mu, lam = 30.0, 30.0
    from mbem_amd.cases import readout_case
    pts, tv, nrm, eps, dens, info = readout_case(args.readout_points)
    ns = tv.shape[0]
    print(f"[readout] points={pts.shape[0]} sources={ns}", flush=True)
    dev = torch.device("cuda")
    src = dv.Sources(torch.from_numpy(tv).to(dev), torch.from_numpy(nrm).to(dev),
                     torch.from_numpy(eps).to(dev))
    X = torch.from_numpy(pts).to(dev)
    D = torch.from_numpy(dens).to(dev)
    out = torch.empty((X.shape[0], 3, 3), dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream()

    chunk = 131072          # points per launch (a few seconds of work each)

    def launches(kind, scale, beta):
        for p0 in range(0, X.shape[0], chunk):
            p1 = min(p0 + chunk, X.shape[0])
            dv.stress(kind, X[p0:p1], src, D, mu, lam, out[p0:p1], scale=scale, beta=beta)

    # small warm launch, then the timed launches with events on the stream
    dv.stress("dd", X[:1024], src, D, mu, lam, out[:1024])
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    times_dd, times_eig = [], []
    for _ in range(max(1, args.steps)):
        ev[0].record(st)
        launches("dd", 1.0, 0.0)
        ev[1].record(st)
        launches("eigen", -1.0, 1.0)
        ev[2].record(st)
        ev[2].synchronize()
        times_dd.append(ev[0].elapsed_time(ev[1]) / 1e3)
        times_eig.append(ev[1].elapsed_time(ev[2]) / 1e3)
        print(f"[readout] dd {times_dd[-1]:.3f} s, eigen {times_eig[-1]:.3f} s", flush=True)
    t_dd, t_eig = float(np.mean(times_dd)), float(np.mean(times_eig))
    pairs = float(pts.shape[0]) * ns
    fl = json.load(open(os.path.join(ROOT, "profiles", "flops_per_pair.json")))
    f_dd, f_eig = fl["dd_stress_pair"]["flops"], fl["eigenstress_pair"]["flops"]
    ex = _executed_flops()
    total = t_dd + t_eig
    # parity and CPU baseline on a bounded sample of the same points
    cpu = None
    err = None
    if not args.no_cpu_baseline:
        from oracle import tri_kernels as ok
        ncpu = args.readout_cpu_points
        sel = np.random.default_rng(7).choice(pts.shape[0], size=ncpu, replace=False)
        t0 = time.perf_counter()
        ref = ok.dd_stress_contract(pts[sel], tv, nrm, eps, dens, mu, lam) \
            - ok.eigenstress_contract(pts[sel], tv, nrm, eps, dens, mu, lam)
        t_cpu = time.perf_counter() - t0
        got = out.cpu().numpy()[sel]
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": t_cpu * pts.shape[0] / ncpu, "unit": "s", "cores": cores, "kind": "port",
               "sample": f"oracle C restatement (OpenMP) on {ncpu} of the {pts.shape[0]} points, "
                         f"all {ns} sources, extrapolated linearly in points "
                         f"({t_cpu:.1f} s measured)"}
        print(f"[readout] oracle on {ncpu} points: {t_cpu:.1f} s, max rel err {err:.2e}",
              flush=True)
    line = {
        "metric": "stress readout time (dd - eigen, FP64, full sum)",
        "value": total, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": 1e3 * total, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic points and densities generated in bench.py",
        "config": {"workload": f"{pts.shape[0]} observation points x {ns} source triangles "
                               f"(N=20,000 box + 676 fault), per-element eps = 0.5 h",
                   "pairs": pairs, "dd_s": t_dd, "eigen_s": t_eig,
                   "pairs_per_s": pairs / total,
                   "flops_per_pair_as_written": {"dd": f_dd, "eigen": f_eig},
                   "algorithmic_tflops_as_written_dd": pairs * f_dd / t_dd / 1e12,
                   "executed_flops_per_pair": {"dd": ex.get("dd"), "eigen": ex.get("eigen")},
                   "max_rel_err_vs_oracle_sample": err},
        "roofline": {"bound": "fp64-valu",
                     "achieved": pairs * ex["dd"] / t_dd / 1e12 if "dd" in ex else None,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": pairs * ex["dd"] / t_dd / 1e12 / FP64_PEAK_TFLOPS if "dd" in ex else None,
                     "traffic": None,
                     "kernel": "k_contract<3> (dd stress; executed FP64 flops per pair from the "
                               "PMC summary fp64_executed_flops.md)",
                     "avg_ms": 1e3 * t_dd},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def learned_ab(system, eps, args):
    """Plain FP64 FGMRES with block-Jacobi alone and with the learned rank-16
    correction (committed weights) on the same assembly: iterations, solve
    time, the network's inference time.  Reported beside the headline; the
    correction is not used by either timed method (it costs iterations)."""
    import torch
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.la.solver import fgmres
    from mbem_amd.learned import WEIGHTS, LearnedPreconditioner, PointNet
    if not os.path.exists(WEIGHTS):
        return None
    asm = AssembledDense(system, eps, "direct", jump="calibrated")
    M = asm.preconditioner(0, args.bj_chunk, grouping=args.grouping)
    LP = LearnedPreconditioner(M, asm.layout, net=PointNet.load(), bf16=True)
    out = {"weights": os.path.relpath(WEIGHTS, ROOT), "used": False,
           "inference_ms": 1e3 * LP.infer_s}
    for label, P in (("block_jacobi", M), ("learned", LP)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, rep = fgmres(asm, asm.b_dev, M=P, rtol=args.rtol, restart=200, maxiter=600)
        torch.cuda.synchronize()
        out[f"iterations_{label}"] = rep.iterations
        out[f"fgmres_s_{label}"] = time.perf_counter() - t0
    print(f"[bench] learned A/B: {out}", flush=True)
    asm.release()
    asm.A_dev = None
    _release(full=True)
    return out


def bench_lu(args):
    """Dense LU of the assembled system (the reference's default solve,
    dense.py:532-571, lu.py:110-161): rocSOLVER getrf + getrs in FP64 on the
    device, n = 60,000 at the default 50 cells.  Prints one JSON line."""
    import torch
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    model, eps, info = box_fault_model(args.n_cells, fault_cells=tuple(args.fault_cells))
    system = generate_system(model)
    n = system.layout.n_unknowns
    how = args.lu_how
    asm = AssembledDense(system, eps, "direct", jump="calibrated", lu=how)
    x_it = asm.solve(rtol=1e-13, precond_max_dense=0, precond_bj_chunk=args.bj_chunk,
                     return_device=True, precond_grouping=args.grouping).clone()
    asm.release()
    times, factor = [], []
    x = None
    for k in range(args.warmup + args.steps):
        asm.release()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = asm.solve(method="lu", return_device=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        fs = asm.report.lu["factor_s"]
        print(f"[lu] {'warmup' if k < args.warmup else 'step'} {k}: {dt:.3f} s (getrf {fs:.3f} s, "
              f"residual {asm.report.residual:.2e}; {asm.report.lu})", flush=True)
        if k >= args.warmup:
            times.append(dt)
            factor.append(fs)
    err = float(torch.linalg.vector_norm(x - x_it) / torch.linalg.vector_norm(x_it))
    flops = 2.0 / 3.0 * n ** 3
    f_s = float(np.mean(factor))
    cpu = None
    if not args.no_cpu_baseline:
        import scipy.linalg
        m = args.lu_cpu_block
        blk = asm.A_dev[:m, :m].cpu().numpy()
        t0 = time.perf_counter()
        scipy.linalg.lu_factor(blk, overwrite_a=True, check_finite=False)
        t_m = time.perf_counter() - t0
        cpu = {"value": t_m * (n / m) ** 3, "unit": "s", "cores": len(os.sched_getaffinity(0)),
               "kind": "port",
               "sample": f"scipy.linalg.lu_factor (LAPACK getrf) of the leading {m} x {m} block "
                         f"of the assembled matrix ({t_m:.2f} s), scaled by (n/m)^3 to n = {n}"}
        print(f"[lu] cpu baseline {cpu['value']:.1f} s", flush=True)
    value = float(np.mean(times))
    single = how != "double"
    peak = FP32_PEAK_TFLOPS if single else FP64_PEAK_TFLOPS
    line = {
        "metric": ("dense LU solve time (N=20k tris, FP32 factors + FP64 refinement)" if single
                   else "dense LU solve time (N=20k tris, FP64)"), "value": value, "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * value,
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64" if single else "f64",
        "data": "synthetic box-with-dipping-fault mesh generated in bench.py",
        "config": {"workload": f"box 40x40x20 km, N={info['box_triangles']} P0 triangles (n={n})",
                   "lu": how, "getrf_s": f_s, "rel_diff_vs_fgmres_1e-13": err,
                   "residual": asm.report.residual,
                   "refinement": {k: v for k, v in asm.report.lu.items()
                                  if k in ("passes", "correction", "left", "fallback")}},
        "roofline": {"bound": "mfma", "achieved": flops / f_s / 1e12, "peak": peak,
                     "unit": "TFLOP/s", "frac": flops / f_s / 1e12 / peak,
                     "traffic": None,
                     "kernel": ("rocSOLVER sgetrf" if single else "rocSOLVER dgetrf")
                               + " (2/3 n^3 flops; factor_s includes the FP32 copy)",
                     "avg_ms": 1e3 * f_s},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-cells", type=int, default=50)
    ap.add_argument("--fault-cells", type=int, nargs=2, default=(13, 26))
    ap.add_argument("--target-err", type=float, default=1e-8)
    ap.add_argument("--rtol", type=float, default=2e-11)
    ap.add_argument("--bj-chunk", type=int, default=1500)
    ap.add_argument("--grouping", choices=("patch", "region"), default="region")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--method", choices=("plain", "mixed", "both"), default="both")
    ap.add_argument("--storage", choices=("bf16", "fp32", "int8"), default="int8")
    ap.add_argument("--inner-rtol", type=float, default=5e-2)
    ap.add_argument("--bj-chunk-mixed", type=int, default=1500)
    ap.add_argument("--grouping-mixed", choices=("patch", "region"), default="region")
    ap.add_argument("--coarse", choices=("none", "const", "rigid", "linear"), default="rigid",
                    help="coarse space of the mixed method's two-level block-Jacobi")
    ap.add_argument("--outer", choices=("fp64", "split"), default="split",
                    help="operator of the mixed method's outer loop with int8 storage: the FP64 "
                         "matrix, or the int8 codes plus their 24-bit rounding remainders")
    ap.add_argument("--other-storage-steps", type=int, default=3,
                    help="steps of the mixed method with the other far-field storage (BF16 "
                         "beside int8), reported beside the headline")
    ap.add_argument("--recycle", type=int, default=90,
                    help="recycled vectors kept across the mixed method's inner solves")
    ap.add_argument("--workload", choices=("solve", "laplace", "readout", "lu"), default="solve",
                    help="solve: the N=20k dense solve (the headline metric); laplace: the "
                         "batched Laplace-state solves; readout: stress readout throughput; "
                         "lu: the dense LU solve")
    ap.add_argument("--lu-cpu-block", type=int, default=6000)
    ap.add_argument("--lu-how", choices=("double", "single", "gpu"), default="double",
                    help="LU flavour of --workload lu (single/gpu: FP32 factors, FP64 refinement)")
    ap.add_argument("--readout-points", type=int, default=1_000_000)
    ap.add_argument("--readout-cpu-points", type=int, default=200)
    ap.add_argument("--laplace-cells", type=int, default=36)
    ap.add_argument("--laplace-fault-cells", type=int, nargs=2, default=(9, 18))
    ap.add_argument("--laplace-rhs", type=int, default=8)
    ap.add_argument("--laplace-rtol", type=float, default=1e-10)
    ap.add_argument("--laplace-restart", type=int, default=30)
    ap.add_argument("--laplace-anchors", type=int, default=4)
    ap.add_argument("--laplace-seq-states", type=int, default=2)
    ap.add_argument("--probe-traffic", type=int, default=0, metavar="N",
                    help="only run the matvec / copy traffic probe on an N x N matrix")
    args = ap.parse_args()
    if args.probe_traffic:
        kernel_probe(args.probe_traffic, 3)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the dense path's sharding is described, not run)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "laplace":
        bench_laplace(args)
        return
    if args.workload == "readout":
        bench_readout(args)
        return
    if args.workload == "lu":
        bench_lu(args)
        return
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
        _release()
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
    far = outer_far = None
    if args.method in ("mixed", "both"):
        # the mixed method's far-field matvec, the mixed step's dominant kernel
        from mbem_amd._lib import call, ptr, stream
        from mbem_amd.la.mixed import MixedOperator
        split = args.storage == "int8" and args.outer == "split"
        op = MixedOperator(asm, storage=args.storage, split=split)
        call("mbem_vec_to_f32", ptr(x_ref), n, ptr(op.xf), stream())
        far_ms = _events_time(lambda: op._far(0, n, n, 1.0, 0.0, y), 10)
        if split:
            # the outer loop's split far-field matvec, timed the same way
            outer_ms = _events_time(
                lambda: call("mbem_gemv_i8e", ptr(op.F), op.ldf, ptr(op.E), op.lde, ptr(op.S),
                             op.nseg, n, n, ptr(x_ref), ptr(y), 1.0, 0.0, stream()), 10)
            outer_far = ("k_gemv_i8e", outer_ms, i8e_bytes(n))
            print(f"[bench] k_gemv_i8e (split outer) {outer_ms:.3f} ms = "
                  f"{i8e_bytes(n) / outer_ms / 1e6:.0f} GB/s", flush=True)
        far_bytes = {"int8": i8_bytes(n), "bf16": 2 * n * n + 12 * n,
                     "fp32": 4 * n * n + 12 * n}[args.storage]
        kname = {"int8": "k_gemv_i8", "bf16": "k_gemv_lo", "fp32": "k_gemv_lo"}[args.storage]
        # what a plain stream reaches at the same byte count: the FP64 matvec
        # over an n x n/8 matrix (the int8 matrix's bytes, rows of the same length)
        nb = n // 8
        B = A[:, :nb].contiguous()
        xb = x_ref[:nb].contiguous()
        ref_ms = _events_time(lambda: gemv(B, xb, y), 10)
        ref_bytes = n * nb * 8 + nb * 8 + n * 8
        del B, xb
        far = (kname, far_ms, far_bytes, ref_ms, ref_bytes, f"k_gemv, {n} x {nb} FP64")
        print(f"[bench] {kname} ({args.storage}) {far_ms:.3f} ms = "
              f"{far_bytes / far_ms / 1e6:.0f} GB/s; k_gemv on {n} x {nb} FP64 {ref_ms:.3f} ms = "
              f"{ref_bytes / ref_ms / 1e6:.0f} GB/s", flush=True)
        del op
    asm = None
    _release(full=True)
    print(f"[bench] gemv {gemv_ms:.3f} ms = {gemv_bytes / gemv_ms / 1e6:.0f} GB/s setup "
          f"{time.perf_counter() - t_setup:.1f} s", flush=True)

    def timed(step_fn, label, steps):
        times, errs, iters, asm_s = [], [], [], []
        asm_k = x = rep = None
        torch.cuda.synchronize()
        for k in range(steps):
            asm_k = x = None
            _release(full=k == 0)
            # Python's cyclic collector is off inside a timed step, as the
            # standard library's timeit does, and runs between steps
            # (_release): a full collection over the solve's objects in the
            # middle of a step stalls the host for tens of milliseconds
            gc.disable()
            try:
                ts = time.perf_counter()
                asm_k, x, rep = step_fn()
                torch.cuda.synchronize()
                times.append(time.perf_counter() - ts)
            finally:
                gc.enable()
            errs.append(float(torch.linalg.vector_norm(x - x_ref)
                              / torch.linalg.vector_norm(x_ref)))
            iters.append(rep.iterations)
            asm_s.append(asm_k.assembly.seconds)
            ph = ", ".join(f"{k_}={v:.3f}" for k_, v in asm_k.timings.items())
            print(f"[bench] {label} step {k}: {times[-1]:.3f} s ({ph}; {rep}; "
                  f"rel err {errs[-1]:.2e})", flush=True)
        summary = getattr(rep, "precond_summary", None)
        a = asm_k.assembly
        pairs = {"t_stored": a.t_pairs_stored, "t_rowsum_only": a.t_pairs_rowsum_only,
                 "u": a.u_pairs, "rhs": a.rhs_pairs}
        del asm_k, a
        return {"times": times, "errs": errs, "iters": iters, "rep": rep, "x": x,
                "precond": summary, "assembly_s": asm_s, "pairs": pairs}

    def step_plain():
        a, x = step()
        return a, x, a.report

    def step_mixed(storage=None):
        from mbem_amd.la.mixed import solve_mixed
        a = AssembledDense(system, eps, "direct", jump="calibrated")
        st = storage or args.storage
        x, rep = solve_mixed(a, rtol=args.rtol, storage=st,
                             inner_rtol=args.inner_rtol, precond_bj_chunk=args.bj_chunk_mixed,
                             precond_grouping=args.grouping_mixed, recycle=args.recycle,
                             coarse=(None if args.coarse == "none" else args.coarse),
                             outer=(args.outer if st == "int8" else "fp64"))
        a.timings.update(mixed_operator_s=rep.operator_build_s,
                         precond_s=rep.precond_build_s, fgmres_s=rep.solve_s)
        return a, x, rep

    results = {}
    if args.method in ("mixed", "both"):
        step_mixed()          # warm the mixed kernels and rocSOLVER's FP32 path
        torch.cuda.synchronize()
    for label, fn in (("plain", step_plain), ("mixed", step_mixed)):
        if args.method in (label, "both"):
            results[label] = timed(fn, label, args.steps)
    other = "bf16" if args.storage != "bf16" else "int8"
    if args.method in ("mixed", "both") and args.other_storage_steps > 0:
        results["mixed_" + other] = timed(lambda: step_mixed(other), f"mixed {other}",
                                          args.other_storage_steps)
    main_label = "mixed" if "mixed" in results else "plain"
    times, errs, iters = (results[main_label][k] for k in ("times", "errs", "iters"))
    total = float(np.sum(times))
    ms_per_step = 1e3 * total / args.steps
    plain_s = float(np.mean(results["plain"]["times"])) if "plain" in results else None
    mixed_s = float(np.mean(results["mixed"]["times"])) if "mixed" in results else None
    other_s = (float(np.mean(results["mixed_" + other]["times"]))
               if "mixed_" + other in results else None)
    pre = (results["plain"]["precond"] or {}) if "plain" in results else {}
    # the error of each method's last solution against the FP64 reference,
    # per quantity: displacement (u-slots), traction (t-slots), and the
    # subtracted stress on the 50 x 50 surface grid, on the top face (z = 0,
    # half the interior limit there) and 4 eps below it
    accuracy = {}
    for label, res in results.items():
        accuracy[label] = quantity_errors(model, system, eps, res["x"], x_ref)
        res["x"] = None
        print(f"[bench] {label} accuracy vs the FP64 reference: {accuracy[label]}", flush=True)
    learned = learned_ab(system, eps, args) if "plain" in results else None
    cpu = None
    if not args.no_cpu_baseline:
        it_plain = results["plain"]["iters"] if "plain" in results else [60]
        cpu = cpu_baseline(model, eps, system, int(np.median(it_plain)))
        print(f"[bench] cpu baseline {cpu['value']:.1f} s on {cpu['cores']} cores", flush=True)
    def roofline(kernel, ms, nbytes, what, ref=None):
        ratio, src_name = _measured_traffic_ratio(kernel)
        out = {"bound": "hbm", "achieved": nbytes / ms / 1e6, "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "frac": nbytes / ms / 1e6 / HBM_PEAK_GBS,
               "traffic": (None if ratio is None else ratio * nbytes),
               "traffic_source": src_name, "kernel": f"{kernel} ({what})",
               "algorithmic_bytes": nbytes, "avg_ms": ms}
        if ref is not None:
            ref_ms, ref_bytes, ref_what = ref
            out["same_bytes_stream"] = {"kernel": ref_what, "avg_ms": ref_ms,
                                        "achieved": ref_bytes / ref_ms / 1e6,
                                        "ratio": (nbytes / ms) / (ref_bytes / ref_ms)}
        return out

    roof_fp64 = roofline("k_gemv", gemv_ms, gemv_bytes, "FGMRES matvec, n x n FP64")
    # the assembly's executed FP64 rate: the pairs each step's assembly ran x
    # the executed flops per pair of the block kernels (PMC counters), over
    # the assembly's device time (HIP events around all of its launches, so
    # the source preparation, the fault right-hand side and the diagonal
    # updates count as time but not as flops)
    ex = _executed_flops()
    asm_rate = None
    if ex.get("t_block") and ex.get("u_block"):
        res = results[main_label]
        pr = res["pairs"]
        flops = (ex["t_block"] * (pr["t_stored"] + pr["t_rowsum_only"]) + ex["u_block"] * pr["u"])
        sec = float(np.median(res["assembly_s"]))
        asm_rate = {"achieved": flops / sec / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": flops / sec / 1e12 / FP64_PEAK_TFLOPS, "assembly_s": sec,
                    "executed_flops": flops, "pairs": pr,
                    "executed_flops_per_pair": {"t": ex["t_block"], "u": ex["u_block"]},
                    "source": "fp64_executed_flops.md (SQ_INSTS_VALU_*_F64 per dispatch)"}
        print(f"[bench] assembly {sec * 1e3:.1f} ms, {flops:.3e} executed FP64 flops: "
              f"{asm_rate['achieved']:.1f} TFLOP/s = {100 * asm_rate['frac']:.0f} % of peak",
              flush=True)
    if main_label == "mixed" and far is not None:
        roof_main = roofline(far[0], far[1], far[2],
                             f"mixed far-field matvec, n x n {args.storage}", ref=far[3:])
    else:
        roof_main = roof_fp64
    mixed_outer = args.outer if args.storage == "int8" else "fp64"
    outer_desc = ("operator: the FP64 matrix" if mixed_outer == "fp64" else
                  "operator: int8 codes + 24-bit rounding remainders, FP64 sums")
    roof_outer = (roofline(outer_far[0], outer_far[1], outer_far[2],
                           "mixed outer matvec, n x n int8 codes + 24-bit remainders, FP64 sums")
                  if main_label == "mixed" and outer_far is not None else None)
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
        "dtype": (f"f64+{args.storage}" if main_label == "mixed" else "f64"),
        "data": "synthetic box-with-dipping-fault mesh generated in bench.py",
        "config": {
            "workload": f"box 40x40x20 km, N={info['box_triangles']} P0 triangles "
                        f"(n={n}), dipping fault {info['fault_triangles']} triangles, "
                        f"eps=0.5h, mu=30 GPa nu=0.25",
            "method": (f"mixed: {args.storage.upper()} far field + FP64 near field on an FP64 "
                       f"outer FGMRES ({outer_desc}), FP32 block-Jacobi"
                       if main_label == "mixed" else
                       "plain FP64: GPU assembly + block-Jacobi FGMRES"),
            "outer_operator": (mixed_outer if main_label == "mixed" else "fp64"),
            "storage": args.storage if main_label == "mixed" else "fp64",
            "rtol": args.rtol, "rel_err_vs_fp64_reference": max(errs),
            "target_rel_err": args.target_err, "iterations": iters,
            "plain_fp64_solve_s": plain_s, "mixed_solve_s": mixed_s,
            "speedup_vs_plain_fp64": (plain_s / mixed_s if plain_s and mixed_s else None),
            "plain_rel_err": (max(results["plain"]["errs"]) if "plain" in results else None),
            f"mixed_{other}_solve_s": other_s,
            f"mixed_{other}_rel_err": (max(results["mixed_" + other]["errs"])
                                       if "mixed_" + other in results else None),
            "accuracy_vs_fp64_reference": accuracy,
            "precond_build_s": pre.get("build_s"), "precond_chunks": pre.get("chunks"),
            "plain_precond": f"block-Jacobi FP64, {args.grouping} grouping, chunk {args.bj_chunk}",
            "mixed_precond": (f"inner FGMRES to {args.inner_rtol:g} on the mixed operator, "
                              f"block-Jacobi FP32, {args.grouping_mixed} grouping, chunk "
                              f"{args.bj_chunk_mixed}, coarse space {args.coarse}, "
                              f"{args.recycle} recycled vectors across inner solves"),
            "mixed_coarse": (results["mixed"]["rep"].coarse if "mixed" in results else None),
            "learned_preconditioner": learned,
        },
        "roofline": roof_main,
        "roofline_fp64_matvec": roof_fp64,
        "roofline_outer_matvec": roof_outer,
        "roofline_assembly": asm_rate,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
