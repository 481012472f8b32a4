"""GPU parity of the batched Laplace-state solve (rheology.py:479-524,
dense.py:650-687, runner.py:357-391 restated as one lockstep batch).

Bars:
  matrix-core GEMM (every tile configuration) and the grouped block-Jacobi
  apply: within 1e-13 (relative max) of a float64 torch reference;
  A(s) = (A0 + g A1) D(s) and b(s) = b0 + g b1: within 1e-11 (relative max
  entry) of the direct assembly at the moduli of s;
  batched solutions: true residual of A(s) = (A0 + g A1) D(s) within 1e-11
  relative at every state; within 1e-9 + 1e-15 lam(s)/mu(s) (relative
  2-norm) of an LU solve of the direct assembly at each state, and of the
  oracle's CPU assembly + LU for two states.  The direct assembly combines the
  reference's t_coeffs, whose lam-slot terms each grow like lam/mu and cancel:
  it loses that many digits toward the fluid limit (the reference's own
  viscoelastic gate, clause a, states the same loss), while the split form
  does not; hence the second term of the bar.
  Graph replays identical to the eager kernel sequence to 1e-13.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TAU = 1.0e19 / 3.0e10          # Maxwell time of a 1e19 Pa s layer with mu = 30 GPa


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def box():
    from mbem_amd import rheology as rh
    from mbem_amd.cases import box_fault_model
    from mbem_amd.kernels import basis as kb
    from mbem_amd.model import generate_system
    from mbem_amd.viscoelastic import LaplaceStates
    model, eps, info = box_fault_model(8, fault_cells=(4, 8))
    # one scalar eps on the relaxing region's closed boundary
    h = float(np.mean([kb.element_sizes(p.mesh).mean() for p in model.regions[0].patches]))
    eps = {p.name: 0.5 * h for p in model.regions[0].patches}
    eps["fault"] = 0.5 * float(kb.element_sizes(model.regions[0].faults[0].mesh).mean())
    system = generate_system(model)
    print(f"case: box n_cells=8, n={system.layout.n_unknowns}, fault 64 triangles", flush=True)
    ls = LaplaceStates(system, eps, {"crust": rh.Maxwell(TAU)}, bj_chunk=300)
    return dict(model=model, eps=eps, system=system, ls=ls, rh=rh)


def _gemm_ref(A0, A1, X, g):
    return A0 @ X + (A1 @ X) * g[None, :]


def test_gemm_g_every_config():
    from mbem_amd import _lib
    from mbem_amd._lib import call, ptr, stream
    from mbem_amd.la import autotune
    torch.manual_seed(1)
    dev = torch.device("cuda")
    m, ld, K = 301, 302, 38
    A0 = torch.randn(m, ld, dtype=torch.float64, device=dev)[:, :m]
    A1 = torch.randn(m, ld, dtype=torch.float64, device=dev)[:, :m]
    X = torch.randn(m, K, dtype=torch.float64, device=dev)
    g = torch.rand(K, dtype=torch.float64, device=dev) + 1.0
    Y0 = torch.randn(m, K, dtype=torch.float64, device=dev)
    ref = _gemm_ref(A0, A1, X, g)
    lib = _lib.load()
    for cfg, c in autotune.configs():
        if not lib.mbem_mgemm_config_fits(cfg, 1):
            # a tile whose two-operator form exceeds the LDS is refused
            with pytest.raises(_lib.MbemError):
                call("mbem_gemm_g", ptr(A0), ptr(A1), ld, m, m, ptr(X), K, K, ptr(g), ptr(X), K,
                     1.0, 0.0, cfg, stream())
            continue
        print(f"gemm_g cfg {cfg} {c}", flush=True)
        Y = torch.empty(m, K, dtype=torch.float64, device=dev)
        call("mbem_gemm_g", ptr(A0), ptr(A1), ld, m, m, ptr(X), K, K, ptr(g), ptr(Y), K, 1.0, 0.0,
             cfg, stream())
        err = float((Y - ref).abs().max() / ref.abs().max())
        assert err < 1e-13, (cfg, err)
        Y2 = Y0.clone()
        call("mbem_gemm_g", ptr(A0), ptr(A1), ld, m, m, ptr(X), K, K, ptr(g), ptr(Y2), K, -2.0, 0.5,
             cfg, stream())
        ref2 = -2.0 * ref + 0.5 * Y0
        assert float((Y2 - ref2).abs().max() / ref2.abs().max()) < 1e-13


def test_grouped_block_jacobi_every_config():
    from mbem_amd.la import autotune
    from mbem_amd.la.batched import GroupedBlockJacobi
    torch.manual_seed(2)
    dev = torch.device("cuda")
    sizes = [37, 64, 5, 100, 33]
    n = sum(sizes)
    K = 38
    A0 = torch.randn(n, n + 1, dtype=torch.float64, device=dev)[:, :n] * 0.1
    A0 += 10.0 * torch.eye(n, dtype=torch.float64, device=dev)
    A1 = torch.randn(n, n + 1, dtype=torch.float64, device=dev)[:, :n] * 0.1
    anchors = [1.2, 1.8]
    ranges = [(0, 20), (20, K)]
    M = GroupedBlockJacobi(A0, A1, sizes, anchors, ranges, K)
    R = torch.randn(n, K, dtype=torch.float64, device=dev)
    ref = torch.zeros_like(R)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for a, (c0, c1) in zip(anchors, ranges):
        for k in range(len(sizes)):
            o0, o1 = int(off[k]), int(off[k + 1])
            B = A0[o0:o1, o0:o1] + a * A1[o0:o1, o0:o1]
            ref[o0:o1, c0:c1] = torch.linalg.solve(B, R[o0:o1, c0:c1])
    for cfg, c in autotune.configs():
        print(f"grouped cfg {cfg} {c}", flush=True)
        M.cfg = cfg
        Z = torch.full_like(R, float("nan"))
        M.apply(R, Z)
        err = float((Z - ref).abs().max() / ref.abs().max())
        assert err < 1e-12, (cfg, err)


def _pair(n, seed, dev):
    """A0, A1 side by side in one n x 2L buffer, as LaplaceStates keeps them."""
    torch.manual_seed(seed)
    L = n + (n & 1)
    AA = torch.zeros(n, 2 * L, dtype=torch.float64, device=dev)
    A0, A1 = AA[:, :n], AA[:, L:L + n]
    A0.copy_(torch.randn(n, n, dtype=torch.float64, device=dev) / math.sqrt(n)
             + 3.0 * torch.eye(n, dtype=torch.float64, device=dev))
    A1.copy_(torch.randn(n, n, dtype=torch.float64, device=dev) / math.sqrt(n))
    return AA, L, A0, A1


def test_library_candidate_matches_kernel():
    from mbem_amd.la.batched import GOperator
    dev = torch.device("cuda")
    n, K = 301, 38
    AA, L, A0, A1 = _pair(n, 11, dev)
    g = torch.rand(K, dtype=torch.float64, device=dev) + 1.0
    X = torch.randn(n, K, dtype=torch.float64, device=dev)
    op = GOperator(A0, A1, g, pair=(AA, L))
    op._prepare()
    Y0 = torch.randn(n, K, dtype=torch.float64, device=dev)
    ref = -0.5 * _gemm_ref(A0, A1, X, g) + 2.0 * Y0
    cands = op.candidates()
    assert 0 in cands and op.LIB in cands and op.LIB + 1 in cands
    for cfg in cands:
        print(f"operator candidate {cfg}: {op.describe(cfg)}", flush=True)
        op.cfg = cfg
        Y = Y0.clone()
        op.apply(X, Y, alpha=-0.5, beta=2.0)
        assert float((Y - ref).abs().max() / ref.abs().max()) < 1e-13, cfg


@pytest.mark.parametrize("use_lib", [False, True])
def test_graph_replays_match_eager(use_lib):
    from mbem_amd.la.batched import BatchedFGMRES, GOperator, GroupedBlockJacobi
    dev = torch.device("cuda")
    n, K = 240, 6
    AA, L, A0, A1 = _pair(n, 3, dev)
    torch.manual_seed(3)
    g = torch.tensor([1.3, 1.3, 1.6, 1.6, 1.9, 1.9], dtype=torch.float64, device=dev)
    sizes = [60, 60, 60, 60]
    M = GroupedBlockJacobi(A0, A1, sizes, [1.4, 1.8], [(0, 4), (4, 6)], K)
    op = GOperator(A0, A1, g, pair=(AA, L))
    if use_lib:
        op.cfg = op.LIB
    Bs = [torch.randn(n, K, dtype=torch.float64, device=dev) for _ in range(2)]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    out = {}
    with torch.cuda.stream(side):
        for graph in (False, True):
            s = BatchedFGMRES(op, M, restart=8, segment=3, graph=graph)
            res = []
            for B in (Bs[0], Bs[1], Bs[0]):
                rep = s.solve(B, rtol=1e-12)
                assert rep.converged.all(), str(rep)
                assert rep.cycles >= 2          # restarts: graphs are replayed
                res.append(s.X.clone())
                side.synchronize()
            s.release()
            out[graph] = res
    torch.cuda.current_stream().wait_stream(side)
    for a, b in zip(out[False], out[True]):
        assert float((a - b).abs().max() / a.abs().max()) < 1e-13
    # and the answer is the solution
    for B, X in zip((Bs[0], Bs[1]), out[True][:2]):
        for c in range(K):
            Bc = A0 + g[c] * A1
            x = torch.linalg.solve(Bc, B[:, c])
            assert float((X[:, c] - x).norm() / x.norm()) < 1e-10


def test_operator_split_matches_direct_assembly(box):
    from mbem_amd.backends.dense import AssembledDense
    ls, rh = box["ls"], box["rh"]
    b0, b1 = ls.rhs_parts([None])
    inv = np.argsort(ls.perm)
    for s in (math.inf, 1.0 / TAU, 0.01 / TAU):
        mu, lam, g = ls.state_moduli(s)
        A_s = ls.operator_at(s).cpu().numpy()
        asm = AssembledDense(box["system"], box["eps"], _materials={"crust": rh.Material(mu, lam)})
        A_d = asm.A
        err = np.abs(A_s - A_d).max() / np.abs(A_d).max()
        b_s = (b0 + g * b1)[:, 0].cpu().numpy()
        err_b = np.abs(b_s - asm.b).max() / np.abs(asm.b).max()
        print(f"s*tau={s * TAU:.3g}: g={g:.6f} mu={mu:.4g} A err {err:.2e} b err {err_b:.2e}",
              flush=True)
        assert err < 1e-11
        assert err_b < 1e-11
        del asm
    assert inv.size == ls.n


def test_batched_states_match_per_state_lu_and_oracle(box):
    from mbem_amd import laplace
    from mbem_amd.backends.dense import AssembledDense
    from oracle import dense as od
    ls, rh, system = box["ls"], box["rh"], box["system"]
    grid = laplace.s_grid(0.01 * TAU, 2e7 * TAU)
    states = rh.laplace_states({"crust": (30.0, 30.0)}, {"crust": rh.Maxwell(TAU)}, grid[::6])
    rng = np.random.default_rng(5)
    fault = box["model"].regions[0].faults[0]
    slip = rng.normal(size=(fault.mesh.n_triangles, 3))
    slip /= np.linalg.norm(slip, axis=1, keepdims=True)
    slips = [None, {"fault": slip}]
    x, rep = ls.solve(states, slips, rtol=1e-12, n_anchors=3)
    print(f"{rep}", flush=True)
    assert rep.batched.converged.all()
    b0, b1 = ls.rhs_parts(slips)
    for i, st in enumerate(states):
        mat = st.materials.get("crust")
        mats = {"crust": mat} if mat is not None else None
        asm = AssembledDense(system, box["eps"], _materials=mats)
        x_lu = asm.solve(method="lu", return_device=True).cpu().numpy()
        e0 = _rel(x[i, 0], x_lu)
        _, _, g = ls.state_moduli(st.s)
        b_r = (b0 + g * b1)[:, 1].cpu().numpy()
        x_r = np.linalg.solve(asm.A, b_r)
        e1 = _rel(x[i, 1], x_r)
        mu, lam, _ = ls.state_moduli(st.s)
        tol = 1e-9 + 1e-15 * lam / mu
        A_s = ls.operator_at(st.s).cpu().numpy()
        b_m = (b0 + g * b1).cpu().numpy()
        r0 = np.linalg.norm(A_s @ x[i, 0] - b_m[:, 0]) / np.linalg.norm(b_m[:, 0])
        r1 = np.linalg.norm(A_s @ x[i, 1] - b_m[:, 1]) / np.linalg.norm(b_m[:, 1])
        print(f"state {st.label} s*tau={st.s * TAU:.3g}: residual {r0:.1e}/{r1:.1e}; vs direct "
              f"LU: model slip {e0:.2e}, random slip {e1:.2e} (bar {tol:.1e})", flush=True)
        assert r0 < 1e-11 and r1 < 1e-11
        assert e0 < tol and e1 < tol
        if i in (0, len(states) - 1):
            A, b, _ = od.assemble(system, box["eps"], materials=mats)
            x_or = od.lu_solve(A, b)
            e_or = _rel(x[i, 0], x_or)
            print(f"  oracle LU err {e_or:.2e} (bar {tol:.1e})", flush=True)
            assert e_or < tol
        del asm


def test_time_map_of_the_batched_stack(box, tmp_path):
    """laplace.py:119-312 on the device stack: the fit's weights applied by
    the library GEMM equal numpy's product to 1e-13, and u(0+) is the elastic
    state (the fit's first row pins it) to 1e-9.  The solved states stored as
    a run directory (runner.py:318-413 layout) read back by ``run_stack`` /
    ``run_fit`` (laplace.py:333-367) bit for bit, with the same fit."""
    from mbem_amd import laplace
    ls, rh = box["ls"], box["rh"]
    grid = laplace.s_grid(0.01 * TAU, 2e7 * TAU)
    states = rh.laplace_states({"crust": (30.0, 30.0)}, {"crust": rh.Maxwell(TAU)}, grid[::3])
    assert states[0].label == "elastic"
    x, rep = ls.solve(states, [None], rtol=1e-12, n_anchors=3)
    assert rep.batched.converged.all()
    stack = x[:, 0]                                      # (n_states, n)
    pr = laplace.Prony([st.s for st in states[1:]])
    t = np.concatenate([[0.0], np.logspace(np.log10(0.1 * TAU), np.log10(1e4 * TAU), 7)])
    W = pr.weights(t)
    got = laplace.apply_weights(W, stack).cpu().numpy()
    ref = W @ stack
    e = float(np.abs(got - ref).max() / np.abs(ref).max())
    e0 = _rel(got[0], stack[0])
    print(f"time map: device vs numpy {e:.1e}; u(0+) vs elastic state {e0:.1e}; fit cond "
          f"{pr.cond:.1e}", flush=True)
    assert e < 1e-13 and e0 < 1e-9
    d = laplace.write_run(tmp_path / "maxwell_run", states, [ls.split(v) for v in stack],
                          {"crust": rh.Maxwell(TAU)},
                          reports=[{"converged": bool(rep.batched.converged.all())}] * len(states))
    s_back, fields = laplace.run_stack(d)
    assert np.array_equal(s_back, np.array([st.s for st in states[1:]]))
    for i, v in enumerate(stack):
        for name, arr in ls.split(v).items():
            assert np.array_equal(fields[name][i], arr)
    assert np.array_equal(laplace.run_fit(d).coefficients(), pr.coefficients())
