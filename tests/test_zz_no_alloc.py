"""The allocation-free multi-dot (mbem_multidot_noalloc) through the C ABI.

Its partial sums live in a fixed device array the library holds, so it
allocates nothing; its results must equal mbem_multidot's (which takes a
caller's work buffer) bit for bit, on two calls in a row with the second
smaller than the first (the array keeps the first call's contents), and a
call that needs more than the array holds is refused with MBEM_ESCRATCH
before anything is launched.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _multidot(lib, V, k, w, h, work, stream):
    from mbem_amd._lib import ptr
    return lib.mbem_multidot(ptr(V), V.stride(0), k, V.shape[1], ptr(w), ptr(h), ptr(work), stream)


def _noalloc(lib, V, k, w, h, stream):
    from mbem_amd._lib import ptr
    return lib.mbem_multidot_noalloc(ptr(V), V.stride(0), k, V.shape[1], ptr(w), ptr(h), stream)


def test_noalloc_multidot_matches_the_workspace_form():
    import torch
    from mbem_amd import _lib
    lib = _lib.load()
    dev = _lib.device()
    st = _lib.stream()
    gen = torch.Generator(device="cpu").manual_seed(7)
    # (rows, columns): a large call, then a smaller one (ragged column count,
    # odd leading dimension so the unaligned path runs too)
    for k, n, ld in ((201, 60_000, 60_000), (37, 20_011, 20_013)):
        V = torch.randn((k, ld), generator=gen, dtype=torch.float64).to(dev)
        w = torch.randn(n, generator=gen, dtype=torch.float64).to(dev)
        Vv = V[:, :n]
        work = torch.empty(int(lib.mbem_multidot_work(k, n)), dtype=torch.float64, device=dev)
        h_ref = torch.empty(k, dtype=torch.float64, device=dev)
        h = torch.full((k,), float("nan"), dtype=torch.float64, device=dev)
        assert _multidot(lib, Vv, k, w, h_ref, work, st) == _lib.MBEM_OK
        assert _noalloc(lib, Vv, k, w, h, st) == _lib.MBEM_OK
        torch.cuda.synchronize()
        assert torch.equal(h, h_ref), (k, n)
        ref = (Vv.cpu() @ w.cpu()).numpy()
        assert np.allclose(h.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_noalloc_multidot_refuses_beyond_its_scratch():
    import torch
    from mbem_amd import _lib
    lib = _lib.load()
    dev = _lib.device()
    n = 512 * 1024                         # 1,024 column chunks of 512
    k = _lib.NOALLOC_SCRATCH_DOUBLES // 1024 + 1
    # a real k x n matrix (4.3 GB), so even a launch would stay in bounds;
    # the call is refused on the host before any launch
    V = torch.zeros((k, n), dtype=torch.float64, device=dev)
    w = torch.zeros(n, dtype=torch.float64, device=dev)
    h = torch.zeros(k, dtype=torch.float64, device=dev)
    null = ctypes.c_void_p(0)
    assert lib.mbem_multidot_noalloc(_lib.ptr(V), n, k, n, _lib.ptr(w), _lib.ptr(h),
                                     _lib.stream()) == _lib.MBEM_ESCRATCH
    # one row fewer fits exactly: ceil(n / 512) * (k - 1) = the array's size
    assert lib.mbem_multidot_noalloc(_lib.ptr(V), n, k - 1, n, _lib.ptr(w), _lib.ptr(h),
                                     _lib.stream()) == _lib.MBEM_OK
    torch.cuda.synchronize()
    assert torch.count_nonzero(h[:k - 1]) == 0
    assert lib.mbem_multidot_noalloc(null, n, -1, n, null, null, null) == _lib.MBEM_EINVAL
    with pytest.raises(_lib.MbemError, match="fixed device scratch"):
        _lib.call("mbem_multidot_noalloc", _lib.ptr(V), n, k, n, _lib.ptr(w), _lib.ptr(h),
                  _lib.stream())
