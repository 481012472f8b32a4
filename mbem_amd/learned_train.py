"""Train and evaluate the learned FGMRES preconditioner on the GPU.

    python -m mbem_amd.learned_train --train 500 --test 50 --out weights.npz

Geometries: ``cases.random_box_fault`` from ``numpy.random.default_rng(seed)``
(fault strike U(0, 180) deg, dip U(20, 90) deg, size U(5, 15) km, box
N = 8 n^2 triangles for even n in ``--cells``).  For each geometry: dense
assembly, block-Jacobi (region grouping, ``--bj-chunk``), residual/correction
pairs from restarted FGMRES cycles (``learned.residual_pairs``), features
and targets on the device.  Training: ``learned.train`` (Adam).  Evaluation
on the held-out geometries: FGMRES iterations and solve time to ``--rtol``
with block-Jacobi alone and with the learned correction.  Prints one JSON
line with the summary.
"""
from __future__ import annotations

import argparse
import gc
import json
import time

import numpy as np
import torch


def build_geometry(rng, cells, bj_chunk, clouds):
    from .backends.dense import AssembledDense
    from .cases import random_box_fault
    from .learned import Cloud
    from .model import generate_system
    model, eps, info = random_box_fault(rng, cells)
    system = generate_system(model)
    asm = AssembledDense(system, eps, "direct", jump="calibrated")
    M = asm.preconditioner(0, bj_chunk, grouping="region")
    key = info["n_cells"]
    if key not in clouds:
        clouds[key] = Cloud(system.layout)
    return asm, M, clouds[key], info


def free(asm):
    """Drop an assembly's device matrix and preconditioner now (the two
    reference each other, so the cyclic collector would free them late)."""
    asm.release()
    asm.A_dev = None
    gc.collect()
    torch.cuda.empty_cache()


def solve_both(asm, M, net, rtol, restart=200, maxiter=600):
    """(iterations, seconds) of FGMRES with block-Jacobi and with the learned
    correction on the same assembled system."""
    from .learned import LearnedPreconditioner
    from .la.solver import fgmres
    out = {}
    LP = LearnedPreconditioner(M, asm.layout, net=net, bf16=True)
    for label, P in (("block_jacobi", M), ("learned", LP)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, rep = fgmres(asm, asm.b_dev, M=P, rtol=rtol, restart=restart, maxiter=maxiter)
        torch.cuda.synchronize()
        out[label] = (rep.iterations, time.perf_counter() - t0, rep.converged, x)
    return out


def main(argv=None):
    from . import _lib
    from .learned import PointNet, evaluate_loss, make_samples, residual_pairs, train
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=40)
    ap.add_argument("--test", type=int, default=8)
    ap.add_argument("--cells", type=int, nargs=2, default=(26, 50))
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--cycles", type=int, default=8)
    ap.add_argument("--restart", type=int, default=10)
    ap.add_argument("--bj-chunk", type=int, default=3000)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--out", default=None)
    ap.add_argument("--init", default=None, help="start from these weights")
    args = ap.parse_args(argv)
    _lib.require_gpu()
    rng = np.random.default_rng(args.seed)
    clouds = {}
    t0 = time.perf_counter()
    samples = []
    for g in range(args.train):
        asm, M, cl, info = build_geometry(rng, args.cells, args.bj_chunk, clouds)
        pairs, rep = residual_pairs(asm, M, cycles=args.cycles, restart=args.restart)
        samples += make_samples(asm, M, cl, pairs)
        print(f"[train-data] geometry {g}: n_cells {info['n_cells']} strike {info['strike']:.0f} "
              f"dip {info['dip']:.0f} size {info['size']:.1f}: {len(pairs)} pairs "
              f"({time.perf_counter() - t0:.0f} s)", flush=True)
        del M, pairs
        free(asm)
    gen_s = time.perf_counter() - t0
    net = PointNet.load(args.init) if args.init else PointNet()
    t0 = time.perf_counter()
    hist = train(net, samples, epochs=args.epochs, lr=args.lr,
                 log=lambda s: print(s, flush=True))
    train_s = time.perf_counter() - t0
    if args.out:
        net.save(args.out, loss_history=np.asarray(hist), seed=args.seed, n_train=args.train)
    # held-out geometries
    res = []
    test_samples = []
    for g in range(args.test):
        asm, M, cl, info = build_geometry(rng, args.cells, args.bj_chunk, clouds)
        pairs, _ = residual_pairs(asm, M, cycles=args.cycles, restart=args.restart)
        test_samples += make_samples(asm, M, cl, pairs)
        out = solve_both(asm, M, net, args.rtol)
        x_bj, x_l = out["block_jacobi"][3], out["learned"][3]
        diff = float(torch.linalg.vector_norm(x_bj - x_l) / torch.linalg.vector_norm(x_bj))
        row = {"n_cells": info["n_cells"], "strike": info["strike"], "dip": info["dip"],
               "size": info["size"],
               "iters_bj": out["block_jacobi"][0], "iters_learned": out["learned"][0],
               "s_bj": out["block_jacobi"][1], "s_learned": out["learned"][1],
               "converged": bool(out["block_jacobi"][2] and out["learned"][2]),
               "rel_diff": diff}
        res.append(row)
        print(f"[test] geometry {g}: n_cells {info['n_cells']}: iterations BJ "
              f"{row['iters_bj']} -> learned {row['iters_learned']}, time {row['s_bj']:.3f} -> "
              f"{row['s_learned']:.3f} s, solutions differ {diff:.1e}", flush=True)
        del M, out
        free(asm)
    test_loss = evaluate_loss(net, test_samples) if test_samples else None
    summary = {
        "train_geometries": args.train, "test_geometries": args.test, "samples": len(samples),
        "epochs": args.epochs, "loss_history": hist, "test_loss": test_loss,
        "data_s": gen_s, "train_s": train_s,
        "mean_iters_bj": float(np.mean([r["iters_bj"] for r in res])) if res else None,
        "mean_iters_learned": float(np.mean([r["iters_learned"] for r in res])) if res else None,
        "mean_s_bj": float(np.mean([r["s_bj"] for r in res])) if res else None,
        "mean_s_learned": float(np.mean([r["s_learned"] for r in res])) if res else None,
        "tests": res,
    }
    print(json.dumps(summary), flush=True)
    return summary


if __name__ == "__main__":
    main()
