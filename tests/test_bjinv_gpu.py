"""The batched FP32 chunk inversion (mbem_bj_gather_f32 / mbem_bj_invert_f32)
against numpy's FP64 inverse.

The chunks are gathered from one FP64 matrix through scattered DOF lists and
are row-permuted, diagonally dominant blocks: well conditioned, yet their
natural diagonal is small, so nearly every step must pivot.  Sizes cover one
row, a partial first panel (3, 31), exact and ragged panels (32, 33, 95), one
to four row slots per thread (513, 1024, 1500, 2048) and a plain Gaussian
block.  Checked: every inverse to FP32 accuracy (1e-5 of its largest entry),
the packed output's padding zero, a singular chunk flagged in info with the
others clean, and the descriptor checks refusing a bad table.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _layout(sizes, B):
    sizes = np.asarray(sizes, dtype=np.int64)
    cptr = np.concatenate([[0], np.cumsum(sizes)])
    ld = ((sizes + 3) // 4) * 4
    iptr = np.concatenate([[0], np.cumsum(sizes * ld)])
    ldw = ((sizes + B - 1) // B) * B
    woff = np.concatenate([[0], np.cumsum(sizes * ldw)])
    soff = np.concatenate([[0], np.cumsum(B * ldw)])
    desc = np.stack([cptr[:-1], sizes, woff[:-1], ldw, iptr[:-1], ld, cptr[:-1], soff[:-1]],
                    axis=1).astype(np.int64)
    return desc, cptr, iptr, ld, woff, soff


def _run(A, chunks):
    import torch
    from mbem_amd import _lib
    from mbem_amd._lib import call, ptr, stream
    lib = _lib.load()
    B = int(lib.mbem_bj_panel_width())
    assert int(lib.mbem_bj_desc_fields()) == 8
    sizes = [c.size for c in chunks]
    desc, cptr, iptr, ld, woff, soff = _layout(sizes, B)
    dev = torch.device("cuda")
    Ad = torch.as_tensor(A, device=dev)
    idx = torch.as_tensor(np.concatenate(chunks).astype(np.int64), device=dev)
    d_dev = torch.as_tensor(desc, device=dev)
    nc, mmax = len(chunks), int(max(sizes))
    W = torch.empty(int(woff[-1]), dtype=torch.float32, device=dev)
    rowmap = torch.empty(int(cptr[-1]), dtype=torch.int32, device=dev)
    piv = torch.empty_like(rowmap)
    saved = torch.empty(int(soff[-1]), dtype=torch.float32, device=dev)
    info = torch.full((nc,), -7, dtype=torch.int32, device=dev)
    out = torch.full((int(iptr[-1]),), 3.0, dtype=torch.float32, device=dev)
    dh = ctypes.c_void_p(desc.ctypes.data)
    call("mbem_bj_gather_f32", ptr(Ad), Ad.stride(0), ptr(idx), idx.numel(), dh, ptr(d_dev), nc,
         mmax, ptr(W), W.numel(), ptr(rowmap), rowmap.numel(), ptr(info), stream())
    call("mbem_bj_invert_f32", dh, ptr(d_dev), nc, mmax, ptr(W), W.numel(), ptr(rowmap),
         ptr(piv), rowmap.numel(), ptr(saved), saved.numel(), ptr(info), ptr(out), out.numel(),
         stream())
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    inv = [o[iptr[c]:iptr[c + 1]].reshape(sizes[c], ld[c]) for c in range(nc)]
    return inv, info.cpu().numpy()


def _permuted_dominant(rng, m):
    """P (4 I + G / sqrt(m)): well conditioned, but the natural diagonal is
    the permuted one's, so a pivot-free elimination would divide by small
    entries."""
    D = 4.0 * np.eye(m) + rng.standard_normal((m, m)) / np.sqrt(m)
    return D[rng.permutation(m)]


def test_batched_inverses_match_numpy():
    rng = np.random.default_rng(11)
    sizes = [1, 3, 31, 32, 33, 95, 513, 1024, 1500, 2048, 200]
    n = sum(sizes) + 500
    A = rng.standard_normal((n, n)) * 1e-3          # off-chunk entries: never read
    perm = rng.permutation(n)
    chunks, o = [], 0
    for c, m in enumerate(sizes):
        dofs = np.sort(perm[o:o + m]) if c % 2 else perm[o:o + m]
        o += m
        blk = _permuted_dominant(rng, m) if c < len(sizes) - 1 else rng.standard_normal((m, m))
        A[np.ix_(dofs, dofs)] = blk
        chunks.append(dofs)
    inv, info = _run(A, chunks)
    assert np.all(info == 0)
    for c, dofs in enumerate(chunks):
        m = dofs.size
        ref = np.linalg.inv(A[np.ix_(dofs, dofs)])
        X = inv[c][:, :m].astype(np.float64)
        err = np.abs(X - ref).max() / np.abs(ref).max()
        cond = np.linalg.cond(A[np.ix_(dofs, dofs)])
        print(f"chunk {c}: m = {m}, cond {cond:.1e}, inverse rel err {err:.2e}")
        assert err < (1e-5 if c < len(sizes) - 1 else 1e-7 * cond * np.sqrt(m))
        assert np.all(inv[c][:, m:] == 0.0)


def test_singular_chunk_is_flagged():
    rng = np.random.default_rng(12)
    sizes = [40, 70, 33]
    n = sum(sizes)
    A = np.zeros((n, n))
    chunks, o = [], 0
    for m in sizes:
        dofs = np.arange(o, o + m)
        A[np.ix_(dofs, dofs)] = _permuted_dominant(rng, m)
        chunks.append(dofs)
        o += m
    A[40 + 17, 40:110] = 0.0                           # chunk 1: a zero row
    inv, info = _run(A, chunks)
    print(f"info {info}")
    assert info[0] == 0 and info[2] == 0 and info[1] > 0
    assert np.all(np.isfinite(inv[1]))
    for c in (0, 2):
        dofs = chunks[c]
        ref = np.linalg.inv(A[np.ix_(dofs, dofs)])
        assert np.abs(inv[c][:, :dofs.size] - ref).max() < 1e-5 * np.abs(ref).max()
