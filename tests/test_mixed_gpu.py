"""Mixed-precision operator and solve on the GPU, on
the 2,048-triangle box with its dipping fault.

  * the near-field blocks are copied exactly (FP64) and the low copy is zero
    there, so (F + N) equals A exactly on the near blocks;
  * the mixed matvec's error against the FP64 matvec is of the storage's
    rounding: below 2e-3 relative (BF16 far field), 2e-7 (FP32) and 2e-3
    (int8 with 64-column segment scales);
  * the int8 codes and scales equal the oracle's (oracle/int8.py) bit for
    bit, the int8 product equals the oracle's dequantized matrix times the
    FP32 vector to FP32 accumulation (1e-5 of the sum of |terms|), and
    holding the near blocks out of A and writing them back leaves A
    bit-identical;
  * the split form (codes plus 24-bit rounding remainders, packed four to
    three words) equals the oracle's bit for bit and its FP64 product the
    oracle's split matrix times x to FP64 rounding (1e-14 of the sum of
    |terms|), within s 2^-25 per entry of A itself;
  * the nested FGMRES (FP64 outer on the exact operator, or on the split
    operator) reaches the FP64 true residual 1e-12 and the oracle's LU
    solution to 1e-9, with and without the recycled space across inner
    solves; the recycled pairs keep
    C^T C = I to 1e-10 and op U = C to 1e-5 (the mixed operator accumulates in
    FP32, so re-applying it to combinations of earlier inputs differs at FP32
    rounding).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    from oracle import dense as od
    model, eps, _ = box_fault_model(16)
    system = generate_system(model)
    asm = DenseBackend().assemble(system, eps)
    A, b, _ = od.assemble(system, eps)
    return dict(asm=asm, A=A, b=b, x_lu=od.lu_solve(A, b))


@pytest.mark.parametrize("storage,tol", [("bf16", 2e-3), ("fp32", 2e-7), ("int8", 2e-3)])
def test_mixed_matvec(c1, storage, tol):
    import torch
    from mbem_amd.la.mixed import MixedOperator
    asm = c1["asm"]
    op = MixedOperator(asm, storage=storage, k_near=2.0)
    vals = op.vals.view(-1, 3, 3).cpu().numpy()
    A = asm.A          # the device's own FP64 matrix: the near copy is exact
    for i in range(0, op.nbr, 97):
        for b in range(op.rowptr_h[i], op.rowptr_h[i + 1]):
            j = op.col_h[b]
            assert np.array_equal(vals[b], A[3 * i:3 * i + 3, 3 * j:3 * j + 3])
    rng = np.random.default_rng(0)
    v = rng.standard_normal(A.shape[0])
    y = op.matvec(torch.as_tensor(v, device=asm.A_dev.device)).cpu().numpy()
    ref = A @ v
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"{storage}: near blocks {op.nnzb}, matvec rel err {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("storage,recycle,outer", [("bf16", 0, "fp64"), ("bf16", 60, "fp64"),
                                                   ("int8", 60, "fp64"), ("int8", 90, "fp64"),
                                                   ("int8", 90, "split")])
def test_mixed_solve(c1, storage, recycle, outer):
    """The last case is the bench's method: int8 far field, 90 recycled
    vectors, rigid coarse space and the outer loop on the split operator
    (int8 codes + 24-bit remainders).  The outer loop then solves the split
    operator's system, not A's: its solution equals the oracle's LU solution
    of the oracle's split operator (far field through oracle/int8.cast_split,
    near blocks exact) to 1e-9, and differs from A's LU solution by the
    representation's floor (entries within s 2^-25 of A's: 2.9e-9 here),
    below the 1e-8 the method is for."""
    from mbem_amd.la.mixed import near_structure, solve_mixed
    from oracle import int8 as oi
    asm = c1["asm"]
    x, rep = solve_mixed(asm, rtol=1e-12, storage=storage, recycle=recycle, outer=outer,
                         coarse="rigid" if outer == "split" else None)
    xh = x.cpu().numpy()
    err = np.linalg.norm(xh - c1["x_lu"]) / np.linalg.norm(c1["x_lu"])
    res64 = np.linalg.norm(c1["b"] - c1["A"] @ xh) / np.linalg.norm(c1["b"])
    print(f"{rep}; rel err vs oracle LU {err:.2e}, oracle FP64 residual {res64:.2e}")
    assert rep.converged and rep.outer.true_relres < 1e-12
    assert rep.outer_operator == outer
    if outer == "fp64":
        assert err < 1e-9 and res64 < 2e-12
        return
    A = c1["A"]
    rowptr, col = near_structure(asm, 2.0)
    near = np.zeros(A.shape, dtype=bool)
    for i in range(rowptr.size - 1):
        for j in col[rowptr[i]:rowptr[i + 1]]:
            near[3 * i:3 * i + 3, 3 * j:3 * j + 3] = True
    codes, scales, resid = oi.cast_split(np.where(near, 0.0, A))
    A_split = np.where(near, A, oi.dequantize_split(codes, scales, resid))
    x_split = np.linalg.solve(A_split, c1["b"])
    err_split = np.linalg.norm(xh - x_split) / np.linalg.norm(x_split)
    print(f"split: rel err vs the oracle split operator's LU {err_split:.2e}")
    assert err_split < 1e-9
    assert err < 1e-8


@pytest.mark.parametrize("m,n", [(37, 2500), (5, 16), (9, 1040), (6, 1030), (45, 5000),
                                 (3, 8208)])
def test_int8_cast_and_product_match_oracle(m, n):
    """Random rows spanning 25 decades, a row 1e-11 smaller, an all-zero
    segment, and column counts that leave a partial last segment, a partial
    last 1024-column block and a tail past the last 16 bytes.  The product
    stages x through LDS 2048 columns at a time: 2500 columns use one tile,
    5000 two (both buffers) with waves past the last row in the second
    workgroup, 8208 eight tiles and a one-vector remainder."""
    import torch
    from mbem_amd._lib import call, load, ptr, stream
    from oracle import int8 as oi
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)) * np.exp(rng.uniform(-20, 5, n))[None, :]
    A[m // 2] *= 1e-11
    A[0, :64] = 0.0
    dev = torch.device("cuda")
    lda = n + 3
    Ad = torch.zeros((m, lda), dtype=torch.float64, device=dev)
    Ad[:, :n] = torch.as_tensor(A, device=dev)
    nseg = int(load().mbem_i8_segments(n))
    assert nseg == oi.segments(n)
    ldf = ((n + 15) // 16) * 16
    F = torch.zeros((m, ldf), dtype=torch.uint8, device=dev)
    S = torch.zeros((m, nseg), dtype=torch.float32, device=dev)
    call("mbem_cast_i8", ptr(Ad), lda, m, n, ptr(F), ldf, ptr(S), nseg, stream())
    codes, scales = oi.cast(A)
    assert np.array_equal(F[:, :n].cpu().numpy(), codes)
    assert np.array_equal(S.cpu().numpy(), scales)
    x = rng.standard_normal(n).astype(np.float32)
    xd = torch.as_tensor(x, device=dev)
    y0 = rng.standard_normal(m)
    y = torch.as_tensor(y0, device=dev)
    call("mbem_gemv_i8", ptr(F), ldf, ptr(S), nseg, m, n, ptr(xd), ptr(y), -0.5, 2.0, stream())
    D = oi.dequantize(codes, scales)
    ref = -0.5 * (D @ x.astype(np.float64)) + 2.0 * y0
    # the kernel sums s (u x) - s (128 x) per 16 columns: the bar is FP32
    # rounding of those terms
    terms = (np.abs(codes.astype(np.float64)) + 128.0) * scales[:, np.arange(n) // 64]
    scale = 0.5 * (terms @ np.abs(x.astype(np.float64))) + 2.0 * np.abs(y0)
    err = np.abs(y.cpu().numpy() - ref) / scale
    print(f"int8 {m}x{n}: product err / |terms| max {err.max():.2e}")
    assert err.max() < 1e-5


@pytest.mark.parametrize("m,n", [(37, 2500), (5, 16), (9, 1040), (6, 1030), (45, 5000),
                                 (3, 8208), (13, 4099), (7, 2050)])
def test_int8_split_cast_and_product_match_oracle(m, n):
    """The split form (codes, scales and packed 24-bit remainders) bit for bit
    against the oracle, and its FP64 product against the oracle's split
    matrix times x: 2500 columns leave one LDS tile (2048) and tail groups of
    four, 5000 two tiles with waves past the last row, 8208 four tiles and
    four tail groups, 4099 two tiles and a 3-column remainder, 2050 one tile
    and a 2-column remainder, 1030 and 16 no tile at all."""
    import torch
    from mbem_amd._lib import call, load, ptr, stream
    from oracle import int8 as oi
    rng = np.random.default_rng(3 * m + n)
    A = rng.standard_normal((m, n)) * np.exp(rng.uniform(-20, 5, n))[None, :]
    A[m // 2] *= 1e-11
    A[0, :64] = 0.0
    dev = torch.device("cuda")
    lda = n + 3
    Ad = torch.zeros((m, lda), dtype=torch.float64, device=dev)
    Ad[:, :n] = torch.as_tensor(A, device=dev)
    nseg = int(load().mbem_i8_segments(n))
    ldf = ((n + 15) // 16) * 16
    nb = int(load().mbem_i8e_resid_bytes(n))
    ldr = nb + 8
    F = torch.zeros((m, ldf), dtype=torch.uint8, device=dev)
    S = torch.zeros((m, nseg), dtype=torch.float32, device=dev)
    E = torch.zeros((m, ldr), dtype=torch.uint8, device=dev)
    call("mbem_cast_i8e", ptr(Ad), lda, m, n, ptr(F), ldf, ptr(S), nseg, ptr(E), ldr, stream())
    codes, scales, resid = oi.cast_split(A)
    assert np.array_equal(F[:, :n].cpu().numpy(), codes)
    assert np.array_equal(S.cpu().numpy(), scales)
    assert np.array_equal(E[:, :nb].cpu().numpy(), oi.pack24(resid))
    x = rng.standard_normal(n)
    xd = torch.as_tensor(x, device=dev)
    y0 = rng.standard_normal(m)
    y = torch.as_tensor(y0, device=dev)
    call("mbem_gemv_i8e", ptr(F), ldf, ptr(E), ldr, ptr(S), nseg, m, n, ptr(xd), ptr(y), -0.5, 2.0,
         stream())
    yh = y.cpu().numpy()
    D = oi.dequantize_split(codes, scales, resid)
    ref = -0.5 * (D @ x) + 2.0 * y0
    # the kernel sums (u x - 128 x) + 2^-32 e x per four columns in FP64: the
    # bar is FP64 rounding of terms up to 256 s |x|
    step = scales[:, np.arange(n) // 64].astype(np.float64)
    scale = 0.5 * ((256.0 * step) @ np.abs(x)) + 2.0 * np.abs(y0)
    err = np.abs(yh - ref) / scale
    print(f"split {m}x{n}: product err / |terms| max {err.max():.2e}")
    assert err.max() < 1e-14
    # and against A itself: the split form is A to within s 2^-25 per entry
    bound = 0.5 * ((step * 2.0 ** -25 + np.abs(A) * 2.0 ** -52) @ np.abs(x)) + 1e-14 * scale
    assert np.all(np.abs(yh - (-0.5 * (A @ x) + 2.0 * y0)) <= bound * 1.01)


def test_near_hold_round_trip(c1):
    import torch
    from mbem_amd._lib import call, ptr, stream
    from mbem_amd.la.mixed import near_structure
    asm = c1["asm"]
    A = asm.A_dev
    before = A.clone()
    rowptr, col = near_structure(asm, 2.0)
    dev = A.device
    rp = torch.as_tensor(rowptr, device=dev)
    cl = torch.as_tensor(col, device=dev)
    nnzb = int(rowptr[-1])
    vals = torch.empty(9 * nnzb, dtype=torch.float64, device=dev)
    call("mbem_near_hold", ptr(A), A.stride(0), rowptr.size - 1, ptr(rp), ptr(cl), nnzb, ptr(vals),
         0, stream())
    held = A.clone()
    call("mbem_near_hold", ptr(A), A.stride(0), rowptr.size - 1, ptr(rp), ptr(cl), nnzb, ptr(vals),
         1, stream())
    assert torch.equal(A, before)
    Bh = before.cpu().numpy()
    Hh = held.cpu().numpy()
    mask = np.zeros(Bh.shape, dtype=bool)
    v = vals.view(-1, 3, 3).cpu().numpy()
    for i in range(rowptr.size - 1):
        for b in range(rowptr[i], rowptr[i + 1]):
            j = col[b]
            mask[3 * i:3 * i + 3, 3 * j:3 * j + 3] = True
            assert np.array_equal(v[b], Bh[3 * i:3 * i + 3, 3 * j:3 * j + 3])
    assert np.all(Hh[mask] == 0.0) and np.array_equal(Hh[~mask], Bh[~mask])


def test_recycled_pairs_stay_consistent(c1):
    """After a few inner solves the recycled (U, C) satisfy op U = C, C^T C = I."""
    import torch
    from mbem_amd.la.mixed import MixedOperator, _RecycledInner
    from mbem_amd.la.preconditioner import BlockGaussSeidel
    asm = c1["asm"]
    op = MixedOperator(asm, storage="bf16")
    M = BlockGaussSeidel(asm, max_dense=0, bj_chunk=1500, grouping="region", inverse_dtype="f32",
                         coupling=op)
    inner = _RecycledInner(op, M, 5e-2, 40, op.n, asm.A_dev.device, kmax=30)
    rng = np.random.default_rng(1)
    for _ in range(4):
        r = torch.as_tensor(rng.standard_normal(op.n), device=asm.A_dev.device)
        z = torch.empty_like(r)
        inner.apply_into(r, z)
    k = inner.k
    assert k > 0
    U, C = inner.U[:k], inner.C[:k]
    OU = torch.stack([op.matvec(U[i].contiguous()) for i in range(k)])
    e1 = float((OU - C).norm() / C.norm())
    e2 = float((C @ C.T - torch.eye(k, dtype=C.dtype, device=C.device)).abs().max())
    print(f"recycled k={k}: |op U - C|/|C| {e1:.2e}, |C^T C - I| {e2:.2e}, "
          f"inner iterations {inner.iterations}")
    assert e1 < 1e-5 and e2 < 1e-10


@pytest.mark.parametrize("storage", ["int8", "bf16"])
def test_mixed_accuracy_per_quantity(c1, storage):
    """The error of the mixed solve at the benchmark's settings (outer rtol
    2e-11, inner 5e-2, 90 recycled vectors) against the oracle's FP64 LU
    solution, stated per quantity: displacement (u-slots), traction
    (t-slots), and the subtracted stress on the 50 x 50 grid over the top
    face (z = 0) and 4 eps below it, each max |error| / max |reference|.
    Bar: the 1e-8 relative error the method is built for."""
    import torch
    from mbem_amd.cases import box_fault_model, surface_grid
    from mbem_amd.evaluate import evaluate_stress
    from mbem_amd.la.mixed import solve_mixed
    asm = c1["asm"]
    layout = asm.layout
    x, rep = solve_mixed(asm, rtol=2e-11, storage=storage, recycle=90)
    xh, xr = x.cpu().numpy(), c1["x_lu"]
    errs = {}
    for kind in ("u", "t"):
        idx = np.concatenate([np.arange(s.offset, s.stop) for s in layout.slots if s.kind == kind])
        errs[kind] = np.abs(xh[idx] - xr[idx]).max() / np.abs(xr[idx]).max()
    model, eps, _ = box_fault_model(16)
    region = asm.system.model.regions[0]
    sol = {s.name: xh[s.offset:s.stop].reshape(-1, 3) for s in layout.slots}
    ref = {s.name: xr[s.offset:s.stop].reshape(-1, 3) for s in layout.slots}
    for tag, z in (("stress_z0", 0.0), ("stress_4eps", -4.0 * eps["top"])):
        pts = surface_grid(50, z=z)
        sa = evaluate_stress(asm.system.model, region, sol, pts, asm.eps, warn_near=False)
        sb = evaluate_stress(asm.system.model, region, ref, pts, asm.eps, warn_near=False)
        errs[tag] = np.abs(sa - sb).max() / np.abs(sb).max()
    torch.cuda.synchronize()
    print(storage, rep, {k: f"{v:.2e}" for k, v in errs.items()})
    assert rep.converged
    for k, v in errs.items():
        assert v < 1e-8, (k, v)
