"""The learned preconditioner's held-out A/B at the scale it was trained for:
the 50 geometries of the training family that follow the 500 training draws
of ``random_box_fault(numpy.random.default_rng(4))`` (box N = 5.4k-20k
triangles, fault strike, dip and size varied), each solved by FGMRES to
relative residual 1e-8 with

* block-Jacobi alone (region grouping, 3,000-DOF chunks, as trained),
* block-Jacobi plus the trained network's rank-16 correction
  (mbem_amd/data/learned_precond.npz, BF16 matrix-core inference),
* block-Jacobi plus the rigid-body coarse space (la/coarse.py),
* block-Jacobi plus the learned coarse space's 16 global columns
  (mbem_amd/data/learned_coarse.npz, learned_coarse.py), alone and beside
  the rigid functions.

Every solve must converge, and the three solutions agree to 2e-5 (a
residual of 1e-8 leaves an error up to the system's conditioning times it:
measured up to 4.7e-6).  The iteration counts
and times to 1e-8 (preconditioner set-up beyond block-Jacobi included) are
printed per geometry and in total.  The coarse space must need fewer
iterations and less time than block-Jacobi alone, on average and on every
geometry (measured: 41.7 against 57.1 iterations, 0.085 s against 0.101 s on
average).  The trained network does not gain (73.4 iterations, 0.152 s):
``test_learned_network_gains`` records that as an expected failure with the
measured ratio; so does ``test_learned_coarse_columns_gain`` for the learned
global columns (measured 60.0 against 57.1 iterations alone, 43.3 against
41.7 beside the rigid functions).
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_TRAIN, N_TEST = 500, 50


def _held_out_params():
    from mbem_amd.cases import random_box_fault_params
    rng = np.random.default_rng(4)
    for _ in range(N_TRAIN):
        random_box_fault_params(rng)
    return [random_box_fault_params(rng) for _ in range(N_TEST)]


@pytest.fixture(scope="module")
def ab():
    import torch
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.cases import random_box_fault
    from mbem_amd.la.coarse import CoarseSpace, TwoLevel
    from mbem_amd.la.solver import fgmres
    from mbem_amd import learned_coarse
    from mbem_amd.learned import LearnedPreconditioner, PointNet
    from mbem_amd.model import generate_system
    net = PointNet.load()
    net_c = PointNet.load(learned_coarse.WEIGHTS)
    labels = ("block_jacobi", "learned", "coarse", "learned_coarse", "rigid_learned_coarse")
    rows = []
    for g, prm in enumerate(_held_out_params()):
        model, eps, info = random_box_fault(None, params=prm)
        system = generate_system(model)
        asm = AssembledDense(system, eps, "direct", jump="calibrated")
        M = asm.preconditioner(0, 3000, grouping="region")
        row = {"n_cells": info["n_cells"], "n": system.layout.n_unknowns}
        xs = {}
        for label in labels:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if label == "block_jacobi":
                P = M
            elif label == "learned":
                P = LearnedPreconditioner(M, asm.layout, net=net, bf16=True)
            elif label == "coarse":
                P = TwoLevel(M, CoarseSpace(asm, M.chunk_dofs, basis="rigid").finish())
            elif label == "learned_coarse":
                P = learned_coarse.preconditioner(asm, M, basis="none", net=net_c)
            else:
                P = learned_coarse.preconditioner(asm, M, basis="rigid", net=net_c)
            x, rep = fgmres(asm, asm.b_dev, M=P, rtol=1e-8, restart=200, maxiter=600)
            torch.cuda.synchronize()
            row[f"it_{label}"] = rep.iterations
            row[f"s_{label}"] = time.perf_counter() - t0
            row[f"ok_{label}"] = rep.converged
            xs[label] = x
        ref = xs["block_jacobi"]
        row["diff"] = max(float(torch.linalg.vector_norm(xs[k] - ref) / torch.linalg.vector_norm(ref))
                          for k in labels[1:])
        rows.append(row)
        its = " / ".join(str(row[f"it_{k}"]) for k in labels)
        secs = " / ".join(f"{row[f's_{k}']:.3f}" for k in labels)
        print(f"[held-out {g}] n_cells {row['n_cells']} (n = {row['n']}): iterations "
              f"{its}; s {secs} ({', '.join(labels)}); solutions differ {row['diff']:.1e}",
              flush=True)
        del asm, M, P, xs, x, ref
    mean = {k: float(np.mean([r[k] for r in rows])) for k in rows[0] if k.startswith(("it_", "s_"))}
    print(f"[held-out] mean over {len(rows)}: {mean}", flush=True)
    return rows, mean


def test_heldout_all_converge_to_one_solution(ab):
    rows, _ = ab
    assert len(rows) == N_TEST
    for r in rows:
        assert all(r[k] for k in r if k.startswith("ok_")), r
        assert r["diff"] < 2e-5, r


def test_coarse_space_gains_on_every_held_out_geometry(ab):
    rows, mean = ab
    assert mean["it_coarse"] < mean["it_block_jacobi"]
    assert mean["s_coarse"] < mean["s_block_jacobi"]
    for r in rows:
        assert r["it_coarse"] < r["it_block_jacobi"], r


@pytest.mark.xfail(reason="the trained network's additive rank-16 correction needs 1.29x the "
                          "iterations (73.4 against 57.1 on average) and 1.50x the time of "
                          "block-Jacobi alone on these geometries at rtol 1e-8: it does not gain",
                   strict=False)
def test_learned_network_gains(ab):
    _, mean = ab
    print(f"learned / block-Jacobi: iterations {mean['it_learned'] / mean['it_block_jacobi']:.2f}, "
          f"time {mean['s_learned'] / mean['s_block_jacobi']:.2f}")
    assert mean["it_learned"] < mean["it_block_jacobi"]
    assert mean["s_learned"] < mean["s_block_jacobi"]


@pytest.mark.xfail(reason="the learned coarse space's 16 global columns need 60.0 iterations "
                          "against 57.1 for block-Jacobi alone and 43.3 against 41.7 beside the "
                          "rigid functions on these geometries; learned_coarse.coarse_bounds shows "
                          "that even the exact slow modes as columns do not gain here",
                   strict=False)
def test_learned_coarse_columns_gain(ab):
    _, mean = ab
    print(f"learned columns / block-Jacobi: iterations "
          f"{mean['it_learned_coarse'] / mean['it_block_jacobi']:.2f}; rigid + learned / rigid: "
          f"{mean['it_rigid_learned_coarse'] / mean['it_coarse']:.2f}")
    assert mean["it_learned_coarse"] < mean["it_block_jacobi"]
    assert mean["it_rigid_learned_coarse"] < mean["it_coarse"]
